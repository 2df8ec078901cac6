// Cull frames through the CadR facade at full size: a transform tree of
// roots x 100 x 100 nodes whose leaves are drawn as instances.  Prints one
// JSON line.
//
//   facade_cull <roots> <frames> [plan|batch]
//
// Scene: 64 Geometries (bounding sphere, 3 LODs as 3 PrimitiveSet records),
// 64 StateSets under one root.  Every leaf gets a random StateSet and
// Geometry; the leaves of each (StateSet, Geometry) pair are the instances of
// one Drawable, whose MatrixList holds their world matrices (4096 Drawables
// for 10M leaves).  Local matrices: random rotation, translation U[-50, 50]^3.
// A frame, on one stream:
//   1. the inner level: world = root x local (one launch; a root's world
//      matrix is its local one);
//   2. the leaf level, propagated straight into the Drawables' MatrixLists
//      (MatrixList::matricesDeviceAddress(), one transform level per list):
//      one launch of a cadrTransformPlan over all lists, its units ordered
//      by parent (`plan`, the default), or cadrPropagateTransformsBatch,
//      1024 lists per launch (`batch`);
//   3. Renderer::recordDrawableProcessing (the reference's gather) and
//      Renderer::recordDrawableCulling for one view (cull items from the
//      MatrixLists through the handle table, cull + LOD + per-StateSet
//      compaction, the reference's draw contract).
// Timing: HIP events around the frame on the stream, and host wall clock from
// beginFrame to the stream's completion; medians over the frames.
#include <CadR/Drawable.h>
#include <CadR/Renderer.h>
#include <cadr_hip.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <random>
#include <stdexcept>
#include <vector>

using namespace CadR;

namespace {

void check(hipError_t e, const char* what)
{
	if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

void checkRc(int rc, const char* what)
{
	if (rc != 0) throw std::runtime_error(std::string(what) + ": " + cadrGetLastErrorMessage());
}

// Random rotation (unit quaternion from three uniforms, Shoemake) and
// translation U[-t, t]^3, column-major.
void rotationTranslation(std::mt19937_64& rng, float t, float* M)
{
	std::uniform_real_distribution<double> u(0.0, 1.0);
	const double u1 = u(rng), u2 = u(rng) * 6.283185307179586, u3 = u(rng) * 6.283185307179586;
	const double a = std::sqrt(1.0 - u1), b = std::sqrt(u1);
	const double x = a * std::sin(u2), y = a * std::cos(u2), z = b * std::sin(u3), w = b * std::cos(u3);
	const double R[9] = {1 - 2 * (y * y + z * z), 2 * (x * y + z * w),     2 * (x * z - y * w),
	                     2 * (x * y - z * w),     1 - 2 * (x * x + z * z), 2 * (y * z + x * w),
	                     2 * (x * z + y * w),     2 * (y * z - x * w),     1 - 2 * (x * x + y * y)};
	for (int c = 0; c < 3; c++) {
		for (int r = 0; r < 3; r++) M[4 * c + r] = float(R[3 * c + r]);
		M[4 * c + 3] = 0.f;
	}
	std::uniform_real_distribution<float> tr(-t, t);
	M[12] = tr(rng); M[13] = tr(rng); M[14] = tr(rng); M[15] = 1.f;
}

uint64_t devAlloc(size_t bytes)
{
	void* p = nullptr;
	check(hipMalloc(&p, std::max<size_t>(bytes, 16)), "hipMalloc");
	return reinterpret_cast<uint64_t>(p);
}

}  // namespace

int main(int argc, char** argv)
{
	if (argc < 3 || (argc > 3 && std::strcmp(argv[3], "plan") != 0 && std::strcmp(argv[3], "batch") != 0)) {
		std::fprintf(stderr, "usage: facade_cull <roots> <frames> [plan|batch]\n");
		return 2;
	}
	const size_t R = std::strtoull(argv[1], nullptr, 10), F1 = 100, F2 = 100;
	const int frames = std::atoi(argv[2]);
	const bool usePlan = argc <= 3 || std::strcmp(argv[3], "plan") == 0;
	const size_t inner = R * F1, leaves = inner * F2;
	const int G = 64, K = 64;
	try {
		Renderer r(0, false);
		std::mt19937_64 rng(0xCAD0000Cull);
		std::vector<std::unique_ptr<Geometry>> geoms;
		for (int g = 0; g < G; g++) {
			geoms.push_back(std::make_unique<Geometry>(r));
			std::vector<uint8_t> v(256), idx(256);
			geoms.back()->uploadVertexData(v.data(), v.size());
			geoms.back()->uploadIndexData(idx.data(), idx.size());
			PrimitiveSet ps[3];
			uint32_t count = 3000, first = 0;
			for (auto& p : ps) {  // LOD 0 finest
				p = {count, first};
				first += count;
				count /= 4;
			}
			geoms.back()->uploadPrimitiveSetData(ps, sizeof(ps));
			std::uniform_real_distribution<float> c(-0.5f, 0.5f), rad(0.5f, 2.f);
			geoms.back()->setBoundingSphere(BoundingSphere{{c(rng), c(rng), c(rng)}, rad(rng)});
			geoms.back()->setNumLods(3);
		}
		std::vector<std::unique_ptr<StateSet>> ss;
		for (int k = 0; k <= K; k++) ss.push_back(std::make_unique<StateSet>(r));
		for (int k = 1; k <= K; k++) ss[0]->childList.append(*ss[k]);

		// tree: roots, inner nodes (parent i / F1), leaves (parent j / F2) grouped by (StateSet, Geometry)
		std::vector<float> rootM(16 * R), innerM(16 * inner), leafM(16 * leaves);
		std::vector<uint32_t> innerP(inner), leafP(leaves);
		for (size_t i = 0; i < R; i++) rotationTranslation(rng, 50.f, &rootM[16 * i]);
		for (size_t i = 0; i < inner; i++) {
			rotationTranslation(rng, 50.f, &innerM[16 * i]);
			innerP[i] = uint32_t(i / F1);
		}
		std::vector<uint32_t> group(leaves);
		std::vector<size_t> groupSize(size_t(K) * G, 0);
		for (size_t i = 0; i < leaves; i++) {
			group[i] = uint32_t(rng() % uint64_t(K)) * G + uint32_t(rng() % uint64_t(G));
			groupSize[group[i]]++;
		}
		std::vector<size_t> groupStart(groupSize.size() + 1, 0);
		for (size_t g = 0; g < groupSize.size(); g++) groupStart[g + 1] = groupStart[g] + groupSize[g];
		{
			std::vector<size_t> at(groupStart.begin(), groupStart.end() - 1);
			for (size_t i = 0; i < leaves; i++) {
				const size_t p = at[group[i]]++;
				rotationTranslation(rng, 50.f, &leafM[16 * p]);
				leafP[p] = uint32_t(i / F2);
			}
		}
		const uint64_t dRoot = devAlloc(rootM.size() * 4), dInnerL = devAlloc(innerM.size() * 4),
		               dInnerW = devAlloc(innerM.size() * 4), dInnerP = devAlloc(innerP.size() * 4),
		               dLeafL = devAlloc(leafM.size() * 4), dLeafP = devAlloc(leafP.size() * 4);
		check(hipMemcpy(reinterpret_cast<void*>(dRoot), rootM.data(), rootM.size() * 4, hipMemcpyHostToDevice), "copy");
		check(hipMemcpy(reinterpret_cast<void*>(dInnerL), innerM.data(), innerM.size() * 4, hipMemcpyHostToDevice), "copy");
		check(hipMemcpy(reinterpret_cast<void*>(dInnerP), innerP.data(), innerP.size() * 4, hipMemcpyHostToDevice), "copy");
		check(hipMemcpy(reinterpret_cast<void*>(dLeafL), leafM.data(), leafM.size() * 4, hipMemcpyHostToDevice), "copy");
		check(hipMemcpy(reinterpret_cast<void*>(dLeafP), leafP.data(), leafP.size() * 4, hipMemcpyHostToDevice), "copy");
		std::vector<float>().swap(leafM);

		// one Drawable per non-empty (StateSet, Geometry) group; its MatrixList holds the group's world matrices
		std::vector<std::unique_ptr<MatrixList>> mls;
		std::vector<std::unique_ptr<Drawable>> drawables;
		std::vector<cadrTransformLevel> leafLevels;
		for (size_t g = 0; g < groupSize.size(); g++) {
			if (groupSize[g] == 0) continue;
			mls.push_back(std::make_unique<MatrixList>(r));
			mls.back()->editNewContent(groupSize[g]);  // the world matrices come from the GPU
			drawables.push_back(std::make_unique<Drawable>(*geoms[g % G], 0u, *mls.back(), *ss[1 + g / G]));
			cadrTransformLevel lv{};
			lv.localMatrices = dLeafL + 64 * groupStart[g];
			lv.parentIndex = dLeafP + 4 * groupStart[g];
			lv.parentWorld = dInnerW;
			lv.parentCount = uint32_t(inner);
			lv.count = groupSize[g];
			leafLevels.push_back(lv);
		}
		r.executeCopyOperations();  // list headers and (placeholder) matrices, once
		for (size_t i = 0; i < leafLevels.size(); i++) leafLevels[i].worldMatrices = mls[i]->matricesDeviceAddress();
		const uint64_t dLevels = devAlloc(leafLevels.size() * sizeof(cadrTransformLevel));
		check(hipMemcpy(reinterpret_cast<void*>(dLevels), leafLevels.data(), leafLevels.size() * sizeof(cadrTransformLevel),
		                hipMemcpyHostToDevice),
		      "copy");
		cadrTransformPlan* plan = nullptr;
		if (usePlan)
			checkRc(cadrTransformPlanCreate(leafLevels.data(), uint32_t(leafLevels.size()), nullptr, &plan),
			        "cadrTransformPlanCreate");
		cadrTransformLevel innerLevel{};
		innerLevel.localMatrices = dInnerL;
		innerLevel.parentIndex = dInnerP;
		innerLevel.parentWorld = dRoot;  // a root's world matrix is its local one
		innerLevel.parentCount = uint32_t(R);
		innerLevel.worldMatrices = dInnerW;
		innerLevel.count = inner;

		cadrView view{};
		const float eye[3] = {0.f, 0.f, 0.f}, dir[3] = {0.f, 0.f, -1.f}, up[3] = {0.f, 1.f, 0.f};
		checkRc(cadrMakeView(eye, dir, up, 1.0471976f, 16.f / 9.f, 0.1f, 2000.f, 1080.f, nullptr, &view), "cadrMakeView");
		hipStream_t stream;
		check(hipStreamCreate(&stream), "hipStreamCreate");
		hipEvent_t e0, e1;
		check(hipEventCreate(&e0), "hipEventCreate");
		check(hipEventCreate(&e1), "hipEventCreate");
		std::vector<double> gpuMs, hostMs;
		uint64_t visible = 0;
		for (int f = 0; f < frames + 2; f++) {
			const auto t0 = std::chrono::steady_clock::now();
			r.beginFrame();
			r.executeCopyOperations();
			check(hipEventRecord(e0, stream), "hipEventRecord");
			checkRc(cadrPropagateTransforms(&innerLevel, stream), "inner level");
			if (usePlan) {
				checkRc(cadrPropagateTransformsPlanned(plan, stream), "leaf levels (plan)");
			} else {
				for (size_t b = 0; b < leafLevels.size(); b += 1024) {
					const uint32_t n = uint32_t(std::min<size_t>(1024, leafLevels.size() - b));
					checkRc(cadrPropagateTransformsBatch(dLevels + b * sizeof(cadrTransformLevel), &leafLevels[b], n,
					                                     stream),
					        "leaf levels");
				}
			}
			r.beginRecording(stream);
			const size_t n = r.prepareSceneRendering(*ss[0]);
			r.recordDrawableProcessing(stream, n);
			r.recordDrawableCulling(stream, &view, 1);
			std::vector<DrawRange> ranges;
			r.recordSceneRendering(stream, *ss[0], ranges);
			r.endRecording(stream);
			check(hipEventRecord(e1, stream), "hipEventRecord");
			r.endFrame();
			check(hipStreamSynchronize(stream), "hipStreamSynchronize");
			const double host = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			float ms = 0.f;
			check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
			if (f >= 2) {  // two warm-up frames (the first builds the cull layout)
				gpuMs.push_back(ms);
				hostMs.push_back(host);
			}
			if (f == frames + 1) {
				std::vector<uint32_t> counts;
				r.downloadCullCounts(counts);
				visible = 0;
				for (uint32_t c : counts) visible += c;
			}
		}
		std::sort(gpuMs.begin(), gpuMs.end());
		std::sort(hostMs.begin(), hostMs.end());
		std::printf("{\"roots\": %zu, \"leaves\": %zu, \"nodes\": %zu, \"drawables\": %zu, \"cull_items\": %zu, "
		            "\"statesets\": %d, \"frames\": %d, \"gpu_ms\": %.4f, \"host_ms\": %.4f, \"visible\": %llu, "
		            "\"handle_level\": %u, \"leaf_propagation\": \"%s\"}\n",
		            R, leaves, R + inner + leaves, drawables.size(), r.numCullItems(), K, frames, gpuMs[gpuMs.size() / 2],
		            hostMs[hostMs.size() / 2], (unsigned long long)visible, r.dataStorage().handleLevel(),
		            usePlan ? "plan" : "batch");
		cadrTransformPlanDestroy(plan);
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		(void)hipStreamDestroy(stream);
		for (uint64_t p : {dRoot, dInnerL, dInnerW, dInnerP, dLeafL, dLeafP, dLevels}) (void)hipFree(reinterpret_cast<void*>(p));
		drawables.clear();
		mls.clear();
		geoms.clear();
		ss.clear();
	} catch (const std::exception& e) {
		std::fprintf(stderr, "facade_cull: %s\n", e.what());
		return 1;
	}
	return 0;
}
