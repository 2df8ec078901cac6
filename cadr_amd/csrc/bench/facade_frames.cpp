// Frame timing of the CadR facade: the same scene and the same edits per
// frame, once with the reference's per-frame full copy of the drawable list
// (StateSet.cpp:233-237 staging memcpy, Renderer.cpp:635-644 copy) and once
// with the resident list and delta uploads.  Prints one JSON line.
//
//   facade_frames <drawables> <frames> <edits per frame>
//
// Scene: 64 StateSets under one root, drawables spread over them in turn,
// 64 geometries with 4 PrimitiveSets each, 4096 MatrixLists of one matrix
// (handle level 2).  An edit re-creates a random drawable in place with
// another geometry, PrimitiveSet and MatrixList (Drawable::create on its own
// StateSet), as an animation or LOD switch would.  A frame is
// beginFrame .. endFrame on one stream, timed on the host from beginFrame to
// the stream's completion; the edits are made before the clock starts.
#include <CadR/Drawable.h>
#include <CadR/Renderer.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <random>
#include <stdexcept>
#include <vector>

using namespace CadR;

namespace {

struct Scene {
	std::vector<std::unique_ptr<Geometry>> geometries;
	std::vector<std::unique_ptr<MatrixList>> matrixLists;
	std::vector<std::unique_ptr<StateSet>> stateSets;
	std::vector<Drawable> drawables;
};

double frameMs(Renderer& r, StateSet& root, hipStream_t stream)
{
	const auto t0 = std::chrono::steady_clock::now();
	r.beginFrame();
	r.executeCopyOperations();
	r.beginRecording(stream);
	const size_t n = r.prepareSceneRendering(root);
	r.recordDrawableProcessing(stream, n);
	std::vector<DrawRange> ranges;
	r.recordSceneRendering(stream, root, ranges);
	r.endRecording(stream);
	r.endFrame();
	if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("hipStreamSynchronize");
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv)
{
	if (argc < 4) {
		std::fprintf(stderr, "usage: facade_frames <drawables> <frames> <edits per frame>\n");
		return 2;
	}
	const size_t N = std::strtoull(argv[1], nullptr, 10);
	const int frames = std::atoi(argv[2]);
	const size_t edits = std::strtoull(argv[3], nullptr, 10);
	try {
		Renderer r(0, false);
		std::mt19937 rng(0xCAD00005u);
		Scene sc;
		const int G = 64, M = 4096, K = 64;
		for (int g = 0; g < G; g++) {
			sc.geometries.push_back(std::make_unique<Geometry>(r));
			std::vector<uint8_t> v(256), idx(256);
			sc.geometries.back()->uploadVertexData(v.data(), v.size());
			sc.geometries.back()->uploadIndexData(idx.data(), idx.size());
			PrimitiveSet ps[4];
			for (auto& p : ps) p = {uint32_t(36 + rng() % 3000), uint32_t(rng() % 100000)};
			sc.geometries.back()->uploadPrimitiveSetData(ps, sizeof(ps));
		}
		const float identity[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
		for (int m = 0; m < M; m++) {
			sc.matrixLists.push_back(std::make_unique<MatrixList>(r));
			sc.matrixLists.back()->setMatrices(identity, 1);
		}
		for (int k = 0; k <= K; k++) sc.stateSets.push_back(std::make_unique<StateSet>(r));
		StateSet& root = *sc.stateSets[0];
		for (int k = 1; k <= K; k++) root.childList.append(*sc.stateSets[k]);
		sc.drawables.reserve(N);
		for (size_t i = 0; i < N; i++)
			sc.drawables.emplace_back(*sc.geometries[rng() % G], uint32_t(8 * (rng() % 4)), *sc.matrixLists[rng() % M],
			                          *sc.stateSets[1 + i % K]);
		hipStream_t stream;
		if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate");

		auto edit = [&]() {
			for (size_t e = 0; e < edits; e++) {
				Drawable& d = sc.drawables[rng() % N];
				d.create(*sc.geometries[rng() % G], uint32_t(8 * (rng() % 4)), *sc.matrixLists[rng() % M], d.stateSet());
			}
		};
		auto run = [&](bool resident, double& medianMs, size_t& bytes) {
			r.setResidentDrawables(resident);
			frameMs(r, root, stream);  // warm-up (and the resident list's first full copy)
			std::vector<double> t;
			size_t sent = 0;
			for (int f = 0; f < frames; f++) {
				edit();
				t.push_back(frameMs(r, root, stream));
				sent += r.lastDrawableUploadBytes();
			}
			std::sort(t.begin(), t.end());
			medianMs = t[t.size() / 2];
			bytes = sent / size_t(frames);
		};
		double fullMs = 0, residentMs = 0;
		size_t fullBytes = 0, residentBytes = 0;
		run(false, fullMs, fullBytes);
		run(true, residentMs, residentBytes);
		std::printf("{\"drawables\": %zu, \"statesets\": %d, \"frames\": %d, \"edits_per_frame\": %zu, "
		            "\"full_copy_ms\": %.4f, \"full_copy_bytes\": %zu, \"resident_ms\": %.4f, \"resident_bytes\": %zu, "
		            "\"handle_level\": %u}\n",
		            N, K, frames, edits, fullMs, fullBytes, residentMs, residentBytes, r.dataStorage().handleLevel());
		(void)hipStreamDestroy(stream);
		sc.drawables.clear();
	} catch (const std::exception& e) {
		std::fprintf(stderr, "facade_frames: %s\n", e.what());
		return 1;
	}
	return 0;
}
