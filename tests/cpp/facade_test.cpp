// Tests of the CadR facade on HIP memory, run by tests/test_facade.py.
//
//   facade_test cpu   device-less Renderer (host memory): handle numbering and
//                     table levels, allocator placement (the checks of the
//                     reference's tests/DataAllocationTest.cpp), the
//                     parent / child lists (tests/ParentChildTest.cpp),
//                     StateSet order, removal and DFS draw ranges
//   facade_test gpu   Renderer on device 0: full frames through
//                     prepareSceneRendering / recordDrawableProcessing at
//                     handle levels 1, 2 and 3, outputs checked against the
//                     handle -> address bookkeeping of the host side; then
//                     frames with scene edits between them (appends,
//                     swap-removals, in-place and cross-StateSet
//                     Drawable::create, StateSet edges), the resident
//                     drawable list checked the same way every frame
//
// Like the reference's tests it prints nothing on success but "OK" and exits
// non-zero with a message on the first failure.
#include <CadR/Drawable.h>
#include <CadR/Exceptions.h>
#include <CadR/Renderer.h>
#include <cadr_hip.h>
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

using namespace CadR;

#define CHECK(cond, msg)                                                                   \
	do {                                                                                   \
		if (!(cond)) throw std::runtime_error(std::string(msg) + " (" #cond ", line " +    \
		                                      std::to_string(__LINE__) + ")");             \
	} while (0)

// Resolve a handle by walking the tables in (host) memory the way
// processDrawables.comp:77-89 does on the device.
static uint64_t walk(const Renderer& r, uint64_t h)
{
	const DataStorage& ds = r.dataStorage();
	auto rd = [](uint64_t a) { return *reinterpret_cast<const uint64_t*>(a); };
	const uint64_t root = ds.handleTableDeviceAddress();
	switch (ds.handleLevel()) {
	case 1: return rd(root + 8 * uint32_t(h));
	case 2: {
		const uint64_t t = rd(root + 8 * uint32_t(h >> 11));
		return t ? rd(t + 8 * (uint32_t(h) & 0x7ff)) : 0;
	}
	default: {
		const uint64_t t2 = rd(root + 8 * uint32_t(h >> 22));
		if (!t2) return 0;
		const uint64_t t3 = rd(t2 + 8 * (uint32_t(h >> 11) & 0x7ff));
		return t3 ? rd(t3 + 8 * (uint32_t(h) & 0x7ff)) : 0;
	}
	}
}

static void testHandleLevels()
{
	Renderer r(false);
	DataStorage& ds = r.dataStorage();
	CHECK(ds.handleLevel() == 0 && ds.handleTableDeviceAddress() == 0, "empty table");
	std::vector<uint64_t> probes = {1, 2, 2047, 2048, 2049, 4095, 4096, 4194303, 4194304, 4194305, 4196351, 4196352,
	                                5000000};
	uint64_t h = 0;
	size_t pi = 0;
	while (h < 5000000) {
		h = ds.createHandle();
		CHECK(h == ds.handleTable().highestHandle(), "handles are sequential");
		const unsigned expLevel = h <= 2047 ? 1 : (h <= 4194303 ? 2 : 3);
		if (pi < probes.size() && h == probes[pi]) {
			CHECK(ds.handleLevel() == expLevel, "level at handle " + std::to_string(h));
			pi++;
		}
		if ((h % 997) == 1 || h == 2047 || h == 2048 || h == 4194303 || h == 4194304)
			ds.setHandle(h, 0x1000000000ull + h * 16);
	}
	CHECK(h == 5000000, "handle count");
	r.executeCopyOperations();
	// every handle set so far resolves through the tables, at the final level
	for (uint64_t x = 1; x <= 5000000; x++) {
		const bool set = (x % 997) == 1 || x == 2047 || x == 2048 || x == 4194303 || x == 4194304;
		const uint64_t exp = set ? 0x1000000000ull + x * 16 : 0;
		if (set || (x % 9973) == 0) CHECK(walk(r, x) == exp, "handle " + std::to_string(x) + " resolves");
	}
	CHECK(walk(r, 0) == 0, "handle 0 resolves to 0");
	// 2442 leaf tables + 2 second-level routing tables + the top table
	CHECK(ds.handleTable().numTables() == (5000000 >> 11) + 1 + 2 + 1, "table count");
}

static void testAllocatorPlacement()
{
	// tests/DataAllocationTest.cpp:55-130: a HandlelessAllocation starts at the
	// beginning of the first block, zero-size allocations have address 0, and
	// the second of two equal allocations of size s lands at the 16/64-byte
	// alignment ladder offset.
	Renderer r(false);
	DataStorage& ds = r.dataStorage();
	HandlelessAllocation a(ds);
	a.alloc(10);
	const uint64_t first = a.deviceAddress();
	a.free();
	a.free();
	r.executeCopyOperations();
	CHECK(ds.liveAllocations() == 0, "storage empty");
	a.alloc(0);
	CHECK(a.deviceAddress() == 0, "zero size allocation has address 0");
	for (size_t s = 0; s < 260; s++) {
		a.alloc(s);
		CHECK(a.deviceAddress() == first || (s == 0 && a.deviceAddress() == 0), "single allocation at the start");
		a.free();
		r.executeCopyOperations();
		CHECK(ds.liveAllocations() == 0, "storage empty after free");
	}
	for (int order = 0; order < 2; order++) {
		for (size_t s = 0; s < 260; s++) {
			HandlelessAllocation a1(ds), a2(ds);
			a1.alloc(s);
			a2.alloc(s);
			const size_t offset = (s == 0) ? 0 : (s <= 16) ? 16 : (s <= 32) ? 32 : (s <= 48) ? 48 : (s <= 64) ? 64
			                                                                                    : (s <= 128) ? 128
			                                                                                                 : (s + 63) & ~size_t(63);
			CHECK(a1.deviceAddress() == first || (s == 0 && a1.deviceAddress() == 0), "first of two at the start");
			CHECK(a2.deviceAddress() == first + offset || (s == 0 && a2.deviceAddress() == 0),
			      "second of two at offset " + std::to_string(offset) + " for size " + std::to_string(s));
			if (order == 0) { a1.free(); a2.free(); } else { a2.free(); a1.free(); }
			r.executeCopyOperations();
			CHECK(ds.liveAllocations() == 0, "storage empty after two frees");
		}
	}
}

// tests/DataAllocationTest.cpp:219-325: for every size s < 260 and every
// count n with n allocations inside the first 64 KiB block, n allocations of
// s bytes lie at the first address + i * offset(s) (the 16/64-byte ladder),
// and every release order (in order, reverse, even then odd, odd then even,
// both reversed, and after a 64 KiB - s allocation pushed the first one to
// the end of the block) leaves the storage empty.
static void testAllocatorSequences()
{
	Renderer r(false);
	DataStorage& ds = r.dataStorage();
	HandlelessAllocation probe(ds);
	probe.alloc(10);
	const uint64_t first = probe.deviceAddress();
	probe.free();
	r.executeCopyOperations();
	auto empty = [&](const char* what, size_t s, size_t n) {
		r.executeCopyOperations();
		CHECK(ds.liveAllocations() == 0,
		      std::string(what) + ": storage not empty (size " + std::to_string(s) + ", n " + std::to_string(n) + ")");
	};
	for (size_t s = 0; s < 260; s++) {
		const size_t offset = (s == 0) ? 0 : (s <= 16) ? 16 : (s <= 32) ? 32 : (s <= 48) ? 48 : (s <= 64) ? 64
		                                                                              : (s <= 128) ? 128
		                                                                                           : (s + 63) & ~size_t(63);
		for (size_t n = 0; n < 1030; n++) {
			if (offset * n >= 65536) continue;
			std::deque<HandlelessAllocation> v;
			auto fill = [&](size_t count) {
				for (size_t i = 0; i < count; i++) v.emplace_back(ds).alloc(s);
			};
			fill(n);
			for (size_t i = 0; i < n; i++)
				CHECK(v[i].deviceAddress() == first + offset * i || (s == 0 && v[i].deviceAddress() == 0),
				      "allocation " + std::to_string(i) + " of " + std::to_string(n) + " of size " +
				          std::to_string(s) + " at first + " + std::to_string(offset * i));
			for (size_t i = 0; i < n; i++) v[i].free();
			v.clear();
			empty("in order", s, n);

			fill(n);
			for (size_t i = n; i > 0;) v[--i].free();
			v.clear();
			empty("reverse", s, n);

			fill(n);
			for (size_t i = 0; i < n; i += 2) v[i].free();
			for (size_t i = 1; i < n; i += 2) v[i].free();
			v.clear();
			empty("even then odd", s, n);

			fill(n);
			for (size_t i = 1; i < n; i += 2) v[i].free();
			for (size_t i = 0; i < n; i += 2) v[i].free();
			v.clear();
			empty("odd then even", s, n);

			fill(n);
			for (int64_t i = int64_t(n) - 2; i >= 0; i -= 2) v[size_t(i)].free();
			for (int64_t i = int64_t(n) - 1; i >= 0; i -= 2) v[size_t(i)].free();
			v.clear();
			empty("every second reversed, from n-2", s, n);

			fill(n);
			for (int64_t i = int64_t(n) - 1; i >= 0; i -= 2) v[size_t(i)].free();
			for (int64_t i = int64_t(n) - 2; i >= 0; i -= 2) v[size_t(i)].free();
			v.clear();
			empty("every second reversed, from n-1", s, n);

			for (int reverse = 0; reverse < 2; reverse++) {
				HandlelessAllocation big(ds);
				big.alloc(65536 - s);
				if (n != 0) fill(1);
				big.free();
				fill(n - (n != 0));
				if (reverse)
					for (size_t i = n; i > 0;) v[--i].free();
				else
					for (size_t i = 0; i < n; i++) v[i].free();
				v.clear();
				empty(reverse ? "after a 64 KiB allocation, reverse" : "after a 64 KiB allocation", s, n);
			}
		}
	}
}

struct Scene {
	std::vector<std::unique_ptr<Geometry>> geometries;
	std::vector<std::unique_ptr<MatrixList>> matrixLists;
	std::vector<std::unique_ptr<DataAllocation>> drawableData;
	std::vector<std::unique_ptr<Drawable>> drawables;
	std::vector<std::unique_ptr<StateSet>> stateSets;
};

static void removeChildOnce(StateSet& parent, StateSet& child)
{
	for (auto it = parent.childList.begin(); it != parent.childList.end(); ++it)
		if (&*it == &child) {
			parent.childList.remove(it);
			return;
		}
	throw std::runtime_error("removeChildOnce: not a child");
}

// tests/ParentChildTest.cpp: the childList / parentList operations, and the
// relation seen from both ends.
static void testParentChildLists()
{
	Renderer r(false);
	StateSet ss1(r), ss2(r), ss3(r);
	CHECK(ss1.childList.size() == 0 && ss1.childList.empty(), "empty child list");
	ChildList::iterator it1 = ss1.childList.append(ss2);
	CHECK(&*it1 == &ss2 && ss1.childList.size() == 1 && &ss1.childList.front() == &ss2 &&
	          &ss1.childList.back() == &ss2,
	      "child appended");
	CHECK(ss2.parentList.size() == 1 && &ss2.parentList.front() == &ss1, "parent seen from the child");
	ss1.childList.remove(it1);
	CHECK(ss1.childList.empty() && ss2.parentList.empty(), "edge removed from both lists");
	ss1.childList.append(ss2);
	ss1.childList.append(ss3);
	ss1.childList.append(ss2);  // a second link to the same child
	size_t k = 0;
	StateSet* expect[3] = {&ss2, &ss3, &ss2};
	for (const StateSet& child : ss1.childList) {
		CHECK(&child == expect[k] && &child.renderer() == &r, "child order is append order");
		k++;
	}
	CHECK(k == 3 && &ss1.childList.back() == &ss2 && ss2.parentList.size() == 2, "three edges");
	ss1.childList.clear();
	CHECK(ss1.childList.empty() && ss2.parentList.empty() && ss3.parentList.empty(), "clear");

	ParentList::iterator it2 = ss2.parentList.append(ss1);
	CHECK(&*it2 == &ss1 && &ss1.childList.front() == &ss2 && &ss2.parentList.back() == &ss1, "parent appended");
	ss2.parentList.append(ss3);
	for (const StateSet& parent : ss2.parentList) parent.renderer();
	ss2.parentList.remove(it2);
	CHECK(ss2.parentList.size() == 1 && ss1.childList.empty() && ss3.childList.size() == 1, "parent edge removed");
	ss2.parentList.clear();
	CHECK(ss2.parentList.empty() && ss3.childList.empty(), "parent clear");
	{
		StateSet temp(r);
		ss1.childList.append(temp);
		temp.childList.append(ss3);
	}
	CHECK(ss1.childList.empty() && ss3.parentList.empty(), "a destroyed StateSet leaves no edges");
}

static void testStateSetOrder()
{
	Renderer r(false);
	Scene sc;
	auto geom = std::make_unique<Geometry>(r);
	PrimitiveSet ps[2] = {{36, 0}, {12, 36}};
	geom->uploadPrimitiveSetData(ps, sizeof(ps));
	for (int i = 0; i < 4; i++) sc.stateSets.push_back(std::make_unique<StateSet>(r));
	StateSet& root = *sc.stateSets[0];
	StateSet& a = *sc.stateSets[1];
	StateSet& b = *sc.stateSets[2];
	StateSet& shared = *sc.stateSets[3];
	root.childList.append(a);
	root.childList.append(b);
	a.childList.append(shared);
	b.childList.append(shared);  // two parents: visited once per path
	auto add = [&](StateSet& ss) {
		sc.matrixLists.push_back(std::make_unique<MatrixList>(r));
		float m[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
		sc.matrixLists.back()->setMatrices(m, 1);
		sc.drawables.push_back(std::make_unique<Drawable>(*geom, 0, *sc.matrixLists.back(), ss));
		return sc.drawables.back().get();
	};
	Drawable* d0 = add(a);
	Drawable* d1 = add(a);
	Drawable* d2 = add(a);
	add(b);
	add(shared);
	add(shared);
	CHECK(root.prepareRecording() == 3 + 1 + 2 + 2, "prepareRecording counts the shared StateSet per path");
	// removal swaps the last drawable into the hole
	const uint64_t ml2 = a.drawableDataList()[2].matrixListHandle;
	d0->destroy();
	CHECK(!d0->isValid() && a.getNumDrawables() == 2, "removed");
	CHECK(a.drawableDataList()[0].matrixListHandle == ml2 && &a.getDrawable(0) == d2 && d2->indexIntoStateSet() == 0,
	      "last drawable moved into the hole");
	CHECK(&a.getDrawable(1) == d1, "untouched drawable keeps its index");
	// move construction keeps the StateSet slot pointing at the new object
	Drawable moved(std::move(*d1));
	CHECK(&a.getDrawable(1) == &moved && !d1->isValid(), "move constructor");
	// DFS ranges: a (2), shared (2), b (1), shared (2)
	std::vector<DrawRange> ranges;
	size_t counter = 0;
	CHECK(root.prepareRecording() == 7, "count after removal");
	std::vector<DrawableGpuData> staging(16);
	root.recordToCommandBuffer(staging.data(), ranges, counter);
	CHECK(counter == 7 && ranges.size() == 4, "four ranges");
	const size_t expBase[4] = {0, 2, 4, 5}, expCount[4] = {2, 2, 1, 2};
	for (int i = 0; i < 4; i++)
		CHECK(ranges[i].base == expBase[i] && ranges[i].count == expCount[i], "range " + std::to_string(i));
	CHECK(ranges[0].stateSet == &a && ranges[1].stateSet == &shared && ranges[2].stateSet == &b &&
	          ranges[3].stateSet == &shared,
	      "range order is DFS");
	CHECK(std::memcmp(&staging[0], &a.drawableDataList()[0], 2 * sizeof(DrawableGpuData)) == 0, "staging content");
	// an empty subtree is skipped; forcing recording keeps it
	StateSet empty(r);
	root.childList.append(empty);
	CHECK(root.prepareRecording() == 7 && empty.skipRecording(), "empty StateSet skipped");
	removeChildOnce(root, empty);
	moved.destroy();
}

// --------------------------------------------------------------------- GPU

static void gpuFrame(int handleLevel, size_t numDrawables, uint32_t seed)
{
	std::printf("gpu frame: target level %d, %zu drawables\n", handleLevel, numDrawables);
	std::fflush(stdout);
	Renderer r(0, false);
	DataStorage& ds = r.dataStorage();
	std::mt19937 rng(seed);
	// burn handles so the table reaches the wanted depth (as a big scene would)
	const uint64_t burn = handleLevel == 1 ? 0 : handleLevel == 2 ? 3000 : 4194304;
	for (uint64_t i = 0; i < burn; i++) ds.createHandle();

	Scene sc;
	const int G = 7;
	std::vector<std::vector<PrimitiveSet>> psData;
	for (int g = 0; g < G; g++) {
		sc.geometries.push_back(std::make_unique<Geometry>(r));
		std::vector<uint8_t> v(64 + 32 * g), idx(48 + 16 * g);
		sc.geometries.back()->uploadVertexData(v.data(), v.size());
		sc.geometries.back()->uploadIndexData(idx.data(), idx.size());
		std::vector<PrimitiveSet> ps(3);
		for (auto& p : ps) p = {uint32_t(rng() % 100000), uint32_t(rng())};
		sc.geometries.back()->uploadPrimitiveSetData(ps.data(), ps.size() * sizeof(PrimitiveSet));
		psData.push_back(ps);
	}
	const int K = 5;
	for (int k = 0; k < K; k++) sc.stateSets.push_back(std::make_unique<StateSet>(r));
	for (int k = 1; k < K; k++) sc.stateSets[k == 4 ? 2 : 0]->childList.append(*sc.stateSets[k]);
	struct Expect {
		int g, p;
		size_t ml;
		int dd;
	};
	std::vector<Expect> expect;
	for (size_t i = 0; i < numDrawables; i++) {
		sc.matrixLists.push_back(std::make_unique<MatrixList>(r));
		const size_t nm = rng() % 4;
		std::vector<float> m(16 * nm, 1.0f);
		sc.matrixLists.back()->setMatrices(m.data(), nm);
		const int g = int(rng() % G), p = int(rng() % 3);
		StateSet& ss = *sc.stateSets[rng() % K];
		int dd = -1;
		if (rng() % 2) {
			sc.drawableData.push_back(std::make_unique<DataAllocation>(ds));
			uint32_t payload[4] = {uint32_t(i), 0, 0, 0};
			sc.drawableData.back()->upload(payload, sizeof(payload));
			dd = int(sc.drawableData.size() - 1);
			sc.drawables.push_back(std::make_unique<Drawable>(*sc.geometries[g], uint32_t(8 * p), *sc.matrixLists.back(),
			                                                  *sc.drawableData.back(), ss));
		} else {
			sc.drawables.push_back(std::make_unique<Drawable>(*sc.geometries[g], uint32_t(8 * p), *sc.matrixLists.back(), ss));
		}
		expect.push_back({g, p, i, dd});
	}
	// remove a few drawables (swap-with-last reorders the StateSets)
	for (size_t i = 0; i < numDrawables / 10; i++) sc.drawables[rng() % numDrawables]->destroy();

	r.setCollectFrameInfo(true);
	hipStream_t stream;
	if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate");
	for (int frame = 0; frame < 2; frame++) {
		r.beginFrame();
		r.executeCopyOperations();
		r.beginRecording(stream);
		const size_t n = r.prepareSceneRendering(*sc.stateSets[0]);
		r.recordDrawableProcessing(stream, n);
		std::vector<DrawRange> ranges;
		r.recordSceneRendering(stream, *sc.stateSets[0], ranges);
		r.endRecording(stream);
		r.endFrame();
		if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("hipStreamSynchronize");
		CHECK(int(ds.handleLevel()) == handleLevel, "handle level " + std::to_string(ds.handleLevel()));
		std::vector<cadrDrawIndirectCommand> ind(n);
		std::vector<cadrDrawablePointers> ptr(n);
		ds.provider().download(ind.data(), r.drawIndirectBufferAddress(), n * sizeof(cadrDrawIndirectCommand));
		ds.provider().download(ptr.data(), r.drawablePointersBufferAddress(), n * sizeof(cadrDrawablePointers));
		// every output slot against the host-side truth of its drawable
		std::unordered_map<const Drawable*, size_t> indexOf;
		for (size_t i = 0; i < sc.drawables.size(); i++) indexOf[sc.drawables[i].get()] = i;
		size_t slot = 0;
		for (const DrawRange& range : ranges) {
			for (size_t j = 0; j < range.count; j++, slot++) {
				const Drawable& d = range.stateSet->getDrawable(j);
				const Expect& e = expect[indexOf.at(&d)];
				const Geometry& g = *sc.geometries[e.g];
				const MatrixList& ml = *sc.matrixLists[e.ml];
				const PrimitiveSet& ps = psData[e.g][e.p];
				uint32_t numMatrices = 0;
				ds.provider().download(&numMatrices, ml.deviceAddress(), 4);
				CHECK(ind[slot].vertexCount == ps.indexCount && ind[slot].firstVertex == ps.startIndex &&
				          ind[slot].instanceCount == numMatrices && ind[slot].firstInstance == 0,
				      "indirect command of slot " + std::to_string(slot));
				CHECK(ptr[slot].vertexDataPtr == g.vertexDataAllocation().deviceAddress() &&
				          ptr[slot].indexDataPtr == g.indexDataAllocation().deviceAddress() &&
				          ptr[slot].matrixListPtr == ml.deviceAddress() &&
				          ptr[slot].drawableDataPtr == (e.dd < 0 ? 0 : sc.drawableData[e.dd]->deviceAddress()),
				      "drawable pointers of slot " + std::to_string(slot));
			}
		}
		CHECK(slot == n, "ranges cover every slot");
		const FrameInfo& fi = r.getFrameInfo();
		CHECK(fi.frameNumber == size_t(frame) && fi.gpuEndExecution >= fi.gpuAfterDrawableProcessingAndBeforeRendering &&
		          fi.gpuAfterDrawableProcessingAndBeforeRendering >= fi.gpuAfterTransfersAndBeforeDrawableProcessing,
		      "frame info timestamps are ordered");
		// second frame: move one MatrixList (new address under the same handle)
		float m[32] = {};
		sc.matrixLists[0]->setMatrices(m, 2);
	}
	(void)hipStreamDestroy(stream);
}

// Frames with edits between them on one renderer: every frame's outputs
// against the host bookkeeping, and the resident list's uploads against what
// the edits touched.
static void gpuResidentFrames(uint32_t seed)
{
	std::printf("gpu resident frames: seed %u\n", seed);
	std::fflush(stdout);
	Renderer r(0, false);
	DataStorage& ds = r.dataStorage();
	std::mt19937 rng(seed);
	for (uint64_t i = 0; i < 3000; i++) ds.createHandle();  // handle level 2

	Scene sc;
	const int G = 5;
	std::vector<std::vector<PrimitiveSet>> psData;
	for (int g = 0; g < G; g++) {
		sc.geometries.push_back(std::make_unique<Geometry>(r));
		std::vector<uint8_t> v(64 + 32 * g), idx(48 + 16 * g);
		sc.geometries.back()->uploadVertexData(v.data(), v.size());
		sc.geometries.back()->uploadIndexData(idx.data(), idx.size());
		std::vector<PrimitiveSet> ps(3);
		for (auto& p : ps) p = {uint32_t(rng() % 100000), uint32_t(rng())};
		sc.geometries.back()->uploadPrimitiveSetData(ps.data(), ps.size() * sizeof(PrimitiveSet));
		psData.push_back(ps);
	}
	// root -> 1, 2; 1 -> 3; 2 -> 3 (StateSet 3 drawn twice); 4 and 5 attached later
	const int K = 6;
	for (int k = 0; k < K; k++) sc.stateSets.push_back(std::make_unique<StateSet>(r));
	StateSet& root = *sc.stateSets[0];
	root.childList.append(*sc.stateSets[1]);
	root.childList.append(*sc.stateSets[2]);
	sc.stateSets[1]->childList.append(*sc.stateSets[3]);
	sc.stateSets[2]->childList.append(*sc.stateSets[3]);
	struct Truth {
		int g, p, ml;
	};
	std::unordered_map<const Drawable*, Truth> truth;
	const int M = 64;
	for (int i = 0; i < M; i++) {
		sc.matrixLists.push_back(std::make_unique<MatrixList>(r));
		std::vector<float> m(16 * (1 + i % 3), 1.0f);
		sc.matrixLists.back()->setMatrices(m.data(), 1 + i % 3);
	}
	auto addDrawable = [&](StateSet& ss) {
		const int g = int(rng() % G), p = int(rng() % 3), ml = int(rng() % M);
		sc.drawables.push_back(std::make_unique<Drawable>(*sc.geometries[g], uint32_t(8 * p), *sc.matrixLists[ml], ss));
		truth[sc.drawables.back().get()] = {g, p, ml};
	};
	for (int i = 0; i < 20000; i++) addDrawable(*sc.stateSets[1 + rng() % 5]);

	hipStream_t stream;
	if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate");
	for (int frame = 0; frame < 14; frame++) {
		size_t touched = 0;  // records the edits of this frame changed (at most)
		bool moved = false;  // the StateSet ranges may move
		if (frame == 2) {
			sc.stateSets[0]->childList.append(*sc.stateSets[4]);
			moved = true;
		}
		if (frame == 5) {
			removeChildOnce(*sc.stateSets[1], *sc.stateSets[3]);
			moved = true;
		}
		if (frame == 7) r.setResidentDrawables(false);
		if (frame == 8) r.setResidentDrawables(true);
		if (frame >= 1 && frame != 4) {
			const int edits = frame % 3 == 0 ? 200 : 20;
			for (int e = 0; e < edits; e++) {
				Drawable& d = *sc.drawables[rng() % sc.drawables.size()];
				const int kind = int(rng() % 4);
				if (kind == 0 && d.isValid()) {
					d.destroy();  // the last record of its StateSet moves into the hole
					moved = true;
				} else if (kind == 1) {
					addDrawable(*sc.stateSets[1 + rng() % 5]);
					moved = true;
				} else if (kind == 2 && d.isValid()) {
					const int g = int(rng() % G), p = int(rng() % 3), ml = int(rng() % M);
					d.create(*sc.geometries[g], uint32_t(8 * p), *sc.matrixLists[ml], d.stateSet());  // in place
					truth[&d] = {g, p, ml};
				} else if (kind == 3 && d.isValid()) {
					const int g = int(rng() % G), p = int(rng() % 3), ml = int(rng() % M);
					d.create(*sc.geometries[g], uint32_t(8 * p), *sc.matrixLists[ml], *sc.stateSets[1 + rng() % 5]);
					truth[&d] = {g, p, ml};
					moved = true;
				}
				touched += 2;
			}
		}
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
		const size_t sent = r.lastDrawableUploadBytes();
		std::printf("  frame %d: %zu drawables, %zu bytes sent\n", frame, n, sent);
		if (frame == 0 || frame == 7 || frame == 8) {
			CHECK(sent == n * sizeof(DrawableGpuData), "full list sent");
		} else if (frame != 2) {  // frame 2 attaches a StateSet the list does not hold yet
			CHECK(sent <= touched * (sizeof(DrawableGpuData) + 4) * 2,
			      std::string("only the changed records sent") + (moved ? " (ranges moved)" : ""));
		}
		if (frame == 4) CHECK(sent == 0, "no edits, nothing sent");
		std::vector<cadrDrawIndirectCommand> ind(n);
		std::vector<cadrDrawablePointers> ptr(n);
		ds.provider().download(ind.data(), r.drawIndirectBufferAddress(), n * sizeof(cadrDrawIndirectCommand));
		ds.provider().download(ptr.data(), r.drawablePointersBufferAddress(), n * sizeof(cadrDrawablePointers));
		size_t slot = 0;
		for (const DrawRange& range : ranges) {
			for (size_t j = 0; j < range.count; j++, slot++) {
				const Drawable& d = range.stateSet->getDrawable(j);
				const Truth& t = truth.at(&d);
				const Geometry& g = *sc.geometries[t.g];
				const MatrixList& ml = *sc.matrixLists[t.ml];
				const PrimitiveSet& ps = psData[t.g][t.p];
				CHECK(ind[slot].vertexCount == ps.indexCount && ind[slot].firstVertex == ps.startIndex &&
				          ind[slot].instanceCount == uint32_t(1 + t.ml % 3),
				      "frame " + std::to_string(frame) + ": indirect command of slot " + std::to_string(slot));
				CHECK(ptr[slot].vertexDataPtr == g.vertexDataAllocation().deviceAddress() &&
				          ptr[slot].indexDataPtr == g.indexDataAllocation().deviceAddress() &&
				          ptr[slot].matrixListPtr == ml.deviceAddress() && ptr[slot].drawableDataPtr == 0,
				      "frame " + std::to_string(frame) + ": drawable pointers of slot " + std::to_string(slot));
			}
		}
		CHECK(slot == n, "ranges cover every slot");
	}
	(void)hipStreamDestroy(stream);
}

int main(int argc, char** argv)
{
	const std::string mode = argc > 1 ? argv[1] : "cpu";
	try {
		if (mode == "cpu") {
			testHandleLevels();
			testAllocatorPlacement();
			testAllocatorSequences();
			testParentChildLists();
			testStateSetOrder();
		} else if (mode == "gpu") {
			gpuFrame(1, 500, 1);
			gpuFrame(2, 20000, 2);
			gpuFrame(3, 20000, 3);
			gpuResidentFrames(4);
			gpuResidentFrames(5);
		} else {
			std::fprintf(stderr, "usage: facade_test cpu|gpu\n");
			return 2;
		}
	} catch (const std::exception& e) {
		std::fprintf(stderr, "FAILED: %s\n", e.what());
		return 1;
	}
	std::printf("OK\n");
	return 0;
}
