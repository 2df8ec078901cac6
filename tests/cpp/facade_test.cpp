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
//                     drawable list checked the same way every frame;
//                     then cull frames (recordDrawableCulling) with one and
//                     with three views, with and without instanced
//                     MatrixLists, some of them written on the GPU by a
//                     transform level, checked against the CPU oracle
//                     (oracle/cadr_oracle.h, linked by this test only)
//
// Like the reference's tests it prints nothing on success but "OK" and exits
// non-zero with a message on the first failure.
#include <CadR/Drawable.h>
#include <CadR/Exceptions.h>
#include <CadR/Renderer.h>
#include <cadr_hip.h>
#include <hip/hip_runtime.h>

#include "../../oracle/cadr_oracle.h"

#include <algorithm>
#include <cmath>
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
// Sustained MatrixList::setMatrices churn (every frame 100 lists get new
// storage of 1..64 matrices, 2 GB over the run) with long-lived allocations
// left behind in every block it touches: freed space inside a block is
// reused (CircularAllocationMemory.h:522-560), blocks that empty go back,
// so the number of memory blocks stays bounded.
static void testAllocatorChurn()
{
	Renderer r(false);
	DataStorage& ds = r.dataStorage();
	std::mt19937 rng(11);
	std::vector<std::unique_ptr<MatrixList>> lists;
	for (int i = 0; i < 100; i++) lists.push_back(std::make_unique<MatrixList>(r));
	std::vector<std::unique_ptr<DataAllocation>> keep;
	std::vector<float> m(16 * 64, 1.f);
	size_t maxBlocks = 0, allocatedBytes = 0;
	for (int frame = 0; frame < 10000; frame++) {
		for (auto& l : lists) {
			const size_t n = 1 + rng() % 64;
			l->setMatrices(m.data(), n);
			allocatedBytes += 64 + 64 * n;
		}
		if (frame % 100 == 0) {  // a long-lived allocation wherever the allocator is at the moment
			keep.push_back(std::make_unique<DataAllocation>(ds));
			uint8_t bytes[256] = {};
			keep.back()->upload(bytes, sizeof(bytes));
		}
		r.executeCopyOperations();
		maxBlocks = std::max(maxBlocks, ds.numBlocks());
		CHECK(ds.liveAllocations() == ds.handleTable().numTables() + lists.size() + keep.size(), "live allocation count");
	}
	CHECK(allocatedBytes > 2000000000ull, "the churn moves gigabytes");
	CHECK(maxBlocks <= 4, "memory blocks stay bounded under churn: " + std::to_string(maxBlocks));
	// everything freed: the storage drains to its first two blocks, reused from their starts
	lists.clear();
	keep.clear();
	CHECK(ds.liveAllocations() == ds.handleTable().numTables(), "storage drained");
	CHECK(ds.numBlocks() <= 2, "emptied blocks outside the allocation order go back");
}

// Renderer::prepareSceneRendering grows the four per-frame buffers to
// 1.2 N slots, at least 128, and only when N exceeds the current capacity
// (Renderer.cpp:461-470): values worked by hand.
static void testBufferGrowth()
{
	Renderer r(false);
	Geometry g(r);
	MatrixList ml(r);
	const float id[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
	ml.setMatrices(id, 1);
	StateSet root(r), child(r);
	root.childList.append(child);
	std::vector<Drawable> ds;
	ds.reserve(2000);
	auto grow = [&](size_t n) {
		while (ds.size() < n) ds.emplace_back(g, 0u, ml, ds.size() % 2 ? child : root);
		return r.prepareSceneRendering(root);
	};
	CHECK(grow(0) == 0 && r.bufferCapacity() == 0, "nothing to draw: no buffers");
	CHECK(grow(1) == 1 && r.bufferCapacity() == 128, "at least 128 slots");
	CHECK(grow(128) == 128 && r.bufferCapacity() == 128, "no growth while N fits");
	CHECK(grow(129) == 129 && r.bufferCapacity() == 154, "1.2 x 129 = 154.8 -> 154 slots");
	CHECK(grow(154) == 154 && r.bufferCapacity() == 154, "no growth at the capacity");
	CHECK(grow(155) == 155 && r.bufferCapacity() == 186, "1.2 x 155 = 186 slots");
	CHECK(grow(1000) == 1000 && r.bufferCapacity() == 1200, "1.2 x 1000 = 1200 slots");
	ds.resize(10);
	CHECK(r.prepareSceneRendering(root) == 10 && r.bufferCapacity() == 1200, "buffers never shrink");
	ds.clear();

	// cull shapes: the one-view shape needs a one-view (or batched) kernel,
	// the batched shape a batched kernel
	r.setCullConfig(cadrCullConfig{1024, 8, 0, 0}, cadrCullConfig{256, 8, 0, 0});
	bool refused = false;
	try {
		r.setCullConfig(cadrCullConfig{1024, 8, 0, 0}, cadrCullConfig{1024, 8, 0, 0});
	} catch (const LogicError&) {
		refused = true;
	}
	CHECK(refused, "a batched cull shape without a batched kernel is refused");
}

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
	hipEvent_t outerBegin, outerEnd;
	if (hipEventCreate(&outerBegin) != hipSuccess || hipEventCreate(&outerEnd) != hipSuccess)
		throw std::runtime_error("hipEventCreate");
	for (int frame = 0; frame < 2; frame++) {
		r.beginFrame();
		r.executeCopyOperations();
		r.beginRecording(stream);
		const size_t n = r.prepareSceneRendering(*sc.stateSets[0]);
		if (hipEventRecord(outerBegin, stream) != hipSuccess) throw std::runtime_error("hipEventRecord");
		r.recordDrawableProcessing(stream, n);
		if (hipEventRecord(outerEnd, stream) != hipSuccess) throw std::runtime_error("hipEventRecord");
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
		CHECK(fi.cpuBeginFrame <= fi.cpuPrepareRecordingBegin && fi.cpuPrepareRecordingBegin <= fi.cpuPrepareRecordingEnd &&
		          fi.cpuPrepareRecordingEnd <= fi.cpuRecordStateSetsBegin &&
		          fi.cpuRecordStateSetsBegin <= fi.cpuRecordStateSetsEnd && fi.cpuRecordStateSetsEnd <= fi.cpuEndFrame &&
		          fi.cpuBeginFrame > 0,
		      "frame info CPU timestamps are taken and ordered");
		// the drawable-processing interval (Renderer.cpp:659-702: ts2 - ts1)
		// lies inside the caller's own events around recordDrawableProcessing
		float outerMs = 0.f;
		if (hipEventElapsedTime(&outerMs, outerBegin, outerEnd) != hipSuccess) throw std::runtime_error("elapsed");
		const uint64_t processing =
		    fi.gpuAfterDrawableProcessingAndBeforeRendering - fi.gpuAfterTransfersAndBeforeDrawableProcessing;
		CHECK(processing > 0 && double(processing) <= double(outerMs) * 1e6 + 20000.0,
		      "drawable processing " + std::to_string(processing) + " ns inside the caller's " +
		          std::to_string(outerMs) + " ms");
		// second frame: move one MatrixList (new address under the same handle)
		float m[32] = {};
		sc.matrixLists[0]->setMatrices(m, 2);
	}
	(void)hipEventDestroy(outerBegin);
	(void)hipEventDestroy(outerEnd);
	(void)hipStreamDestroy(stream);
}

// A frame with no drawables: recordDrawableProcessing dispatches nothing and
// records the two timestamps back to back (Renderer.cpp:600-620), so the
// drawable-processing interval is (near) zero and the frame's timestamps stay
// ordered.
static void gpuEmptyFrame()
{
	std::printf("gpu empty frame\n");
	std::fflush(stdout);
	Renderer r(0, false);
	StateSet root(r);
	r.setCollectFrameInfo(true);
	hipStream_t stream;
	if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate");
	r.beginFrame();
	r.executeCopyOperations();
	r.beginRecording(stream);
	const size_t n = r.prepareSceneRendering(root);
	CHECK(n == 0, "an empty scene counts no drawables");
	r.recordDrawableProcessing(stream, n);
	std::vector<DrawRange> ranges;
	r.recordSceneRendering(stream, root, ranges);
	r.endRecording(stream);
	r.endFrame();
	if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("hipStreamSynchronize");
	const FrameInfo& fi = r.getFrameInfo();
	const uint64_t processing =
	    fi.gpuAfterDrawableProcessingAndBeforeRendering - fi.gpuAfterTransfersAndBeforeDrawableProcessing;
	CHECK(fi.gpuAfterDrawableProcessingAndBeforeRendering >= fi.gpuAfterTransfersAndBeforeDrawableProcessing &&
	          fi.gpuEndExecution >= fi.gpuAfterDrawableProcessingAndBeforeRendering && processing < 200000,
	      "empty frame: processing interval " + std::to_string(processing) + " ns");
	CHECK(ranges.empty() || ranges[0].count == 0, "an empty scene draws nothing");
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

// ------------------------------------------------------------ cull frames

// A column-major rotation (axis-angle) x uniform scale + translation.
static void randomMatrix(std::mt19937& rng, float range, float* M)
{
	std::uniform_real_distribution<float> u(-1.f, 1.f), sc(0.5f, 2.f), ang(0.f, 6.2831853f);
	float ax = u(rng), ay = u(rng), az = u(rng);
	const float len = std::sqrt(ax * ax + ay * ay + az * az) + 1e-6f;
	ax /= len; ay /= len; az /= len;
	const float a = ang(rng), c = std::cos(a), s = std::sin(a), t = 1.f - c, k = sc(rng);
	const float R[9] = {t * ax * ax + c,      t * ax * ay + s * az, t * ax * az - s * ay,
	                    t * ax * ay - s * az, t * ay * ay + c,      t * ay * az + s * ax,
	                    t * ax * az + s * ay, t * ay * az - s * ax, t * az * az + c};
	for (int col = 0; col < 3; col++) {
		for (int row = 0; row < 3; row++) M[4 * col + row] = k * R[3 * col + row];
		M[4 * col + 3] = 0.f;
	}
	M[12] = range * u(rng); M[13] = range * u(rng); M[14] = range * u(rng); M[15] = 1.f;
}

// Cull frames through the facade against the oracle: drawables with one
// MatrixList each (one cull item per drawable) or, with `instancing`, 0 to 3
// matrices (one item per matrix, some lists written on the GPU by a
// transform level); StateSets nested and one reached through two parents;
// a Geometry destroyed while its drawables remain; views from cadrMakeView.
static void gpuCullFrames(uint32_t seed, bool instancing, size_t numDrawables, bool shared = false)
{
	std::printf("gpu cull frames: seed %u, instancing %d, %zu drawables, shared buffers %d\n", seed, int(instancing),
	            numDrawables, int(shared));
	std::fflush(stdout);
	Renderer r(0, false);
	r.setCullBuffersShared(shared);
	uint64_t generation = r.cullBufferGeneration();
	DataStorage& ds = r.dataStorage();
	std::mt19937 rng(seed);
	if (seed % 2) for (int i = 0; i < 3000; i++) ds.createHandle();  // handle level 2

	const int G = 6;
	std::vector<std::unique_ptr<Geometry>> geoms;
	std::vector<std::vector<PrimitiveSet>> psData;
	for (int g = 0; g <= G; g++) {  // geometry G is destroyed before the frames
		geoms.push_back(std::make_unique<Geometry>(r));
		std::vector<uint8_t> v(64), idx(64);
		geoms.back()->uploadVertexData(v.data(), v.size());
		geoms.back()->uploadIndexData(idx.data(), idx.size());
		std::vector<PrimitiveSet> ps(5);  // LODs at offset 0 or 8, up to 4 of them
		for (auto& p : ps) p = {uint32_t(36 + rng() % 3000), uint32_t(rng() % 1000000)};
		geoms.back()->uploadPrimitiveSetData(ps.data(), ps.size() * sizeof(PrimitiveSet));
		psData.push_back(ps);
		std::uniform_real_distribution<float> u(-1.f, 1.f), rad(0.5f, 4.f);
		geoms.back()->setBoundingSphere(BoundingSphere{{u(rng), u(rng), u(rng)}, rad(rng)});
		geoms.back()->setNumLods(1 + uint32_t(g % 4));
	}
	const int K = 6;
	std::vector<std::unique_ptr<StateSet>> ss;
	for (int k = 0; k < K; k++) ss.push_back(std::make_unique<StateSet>(r));
	ss[0]->childList.append(*ss[1]);
	ss[0]->childList.append(*ss[2]);
	ss[1]->childList.append(*ss[3]);
	ss[2]->childList.append(*ss[3]);  // StateSet 3 is drawn twice, once per path
	ss[2]->childList.append(*ss[4]);
	ss[0]->childList.append(*ss[5]);

	struct DInfo { int g; uint32_t off; size_t ml; };
	std::vector<std::unique_ptr<MatrixList>> mls;
	std::vector<std::vector<float>> mlHost;
	std::vector<std::unique_ptr<Drawable>> drawables;
	std::vector<DInfo> info;
	for (size_t i = 0; i < numDrawables; i++) {
		const size_t nm = instancing ? rng() % 4 : 1;
		std::vector<float> m(16 * nm);
		for (size_t t = 0; t < nm; t++) randomMatrix(rng, 300.f, &m[16 * t]);
		mls.push_back(std::make_unique<MatrixList>(r));
		mls.back()->setMatrices(m.data(), nm);
		mlHost.push_back(m);
		const int g = (i % 50 == 7) ? G : int(rng() % G);
		const uint32_t off = 8u * uint32_t(rng() % 2);
		drawables.push_back(std::make_unique<Drawable>(*geoms[g], off, *mls.back(), *ss[rng() % K]));
		info.push_back({g, off, mls.size() - 1});
	}
	for (size_t i = 0; i < numDrawables / 20; i++) drawables[rng() % numDrawables]->destroy();
	// the mesh table the GPU should build, before geometry G goes away
	std::vector<cadrMesh> expMeshes(r.numMeshes());
	for (int g = 0; g <= G; g++) {
		for (uint32_t off : {0u, 8u}) {
			const uint32_t m = r.meshIndexOf(*geoms[g], off);
			if (m >= expMeshes.size()) expMeshes.resize(m + 1);
			cadrMesh& e = expMeshes[m];
			e = cadrMesh{};
			if (g == G) {
				e.radius = -INFINITY;
				e.lodCount = 1;
				continue;
			}
			const BoundingSphere& b = geoms[g]->boundingSphere();
			e.center[0] = b.center[0]; e.center[1] = b.center[1]; e.center[2] = b.center[2]; e.radius = b.radius;
			e.lodCount = std::min<uint32_t>(geoms[g]->numLods(), (40 - off) / 8);
			for (uint32_t k = 0; k < e.lodCount; k++) {
				e.indexCount[k] = psData[g][off / 8 + k].indexCount;
				e.firstIndex[k] = psData[g][off / 8 + k].startIndex;
			}
		}
	}
	geoms[G].reset();
	std::unordered_map<const Drawable*, size_t> indexOf;
	for (size_t i = 0; i < drawables.size(); i++) indexOf[drawables[i].get()] = i;

	// with instancing, the first lists with matrices get world = parent x local from a GPU transform level
	std::vector<void*> deviceArrays;
	auto devUpload = [&](const void* p, size_t bytes) {
		void* d = nullptr;
		if (hipMalloc(&d, std::max<size_t>(bytes, 16)) != hipSuccess) throw std::runtime_error("hipMalloc");
		if (bytes && hipMemcpy(d, p, bytes, hipMemcpyHostToDevice) != hipSuccess) throw std::runtime_error("hipMemcpy");
		deviceArrays.push_back(d);
		return reinterpret_cast<uint64_t>(d);
	};
	std::vector<float> roots(16 * 4);
	for (int i = 0; i < 4; i++) randomMatrix(rng, 100.f, &roots[16 * i]);
	const uint64_t rootsDev = devUpload(roots.data(), roots.size() * 4);
	struct GpuList { size_t ml; uint64_t local, parent; };
	std::vector<GpuList> gpuLists;
	if (instancing) {
		for (size_t i = 0; i < mls.size() && gpuLists.size() < 16; i++) {
			const size_t n = mls[i]->numMatrices();
			if (n == 0) continue;
			std::vector<float> local(16 * n);
			std::vector<uint32_t> parent(n);
			for (size_t t = 0; t < n; t++) {
				randomMatrix(rng, 50.f, &local[16 * t]);
				parent[t] = uint32_t(rng() % 4);
				oracle_mat4_mul(&roots[16 * parent[t]], &local[16 * t], &mlHost[i][16 * t]);
			}
			gpuLists.push_back({i, devUpload(local.data(), local.size() * 4), devUpload(parent.data(), parent.size() * 4)});
		}
	}

	std::vector<cadrView> views(3);
	const float eyes[3][3] = {{0.f, 0.f, 0.f}, {350.f, 50.f, -200.f}, {-100.f, 300.f, 400.f}};
	const float dirs[3][3] = {{0.f, 0.f, -1.f}, {-1.f, -0.1f, 0.6f}, {0.2f, -0.7f, -1.f}};
	const float up[3] = {0.f, 1.f, 0.f};
	const float fovs[3] = {1.0472f, 1.2f, 0.8f};
	for (int v = 0; v < 3; v++)
		CHECK(cadrMakeView(eyes[v], dirs[v], up, fovs[v], 16.f / 9.f, 0.1f, 2000.f, 1080.f, nullptr, &views[v]) == 0,
		      "cadrMakeView");

	hipStream_t stream;
	if (hipStreamCreate(&stream) != hipSuccess) throw std::runtime_error("hipStreamCreate");
	for (int frame = 0; frame < 3; frame++) {
		const size_t V = frame == 0 ? 1 : 3;
		r.setCullDrawablePointers(frame > 0);
		if (frame == 2) {
			// edits: a list changes size, a drawable moves StateSet, one is removed
			std::vector<float> m(16 * 2);
			randomMatrix(rng, 300.f, &m[0]);
			randomMatrix(rng, 300.f, &m[16]);
			const size_t target = instancing ? 3 : 5;
			const size_t nm = instancing ? 2 : 1;
			mls[target]->setMatrices(m.data(), nm);
			mlHost[target].assign(m.begin(), m.begin() + 16 * nm);
			for (GpuList& gl : gpuLists) if (gl.ml == target) gl.ml = size_t(-1);
			for (size_t i = 0; i < drawables.size(); i++)
				if (drawables[i]->isValid() && info[i].g != G) {
					drawables[i]->create(*geoms[info[i].g], info[i].off, *mls[info[i].ml], *ss[(K - 1) - (i % 2)]);
					break;
				}
			for (size_t i = drawables.size(); i-- > 0;)
				if (drawables[i]->isValid()) { drawables[i]->destroy(); break; }
		}
		r.beginFrame();
		r.executeCopyOperations();
		for (const GpuList& gl : gpuLists) {
			if (gl.ml == size_t(-1)) continue;
			cadrTransformLevel lv{};
			lv.localMatrices = gl.local;
			lv.parentIndex = gl.parent;
			lv.parentWorld = rootsDev;
			lv.parentCount = 4;
			lv.worldMatrices = mls[gl.ml]->matricesDeviceAddress();
			lv.count = mls[gl.ml]->numMatrices();
			CHECK(cadrPropagateTransforms(&lv, stream) == 0, "cadrPropagateTransforms into a MatrixList");
		}
		r.beginRecording(stream);
		const size_t n = r.prepareSceneRendering(*ss[0]);
		r.recordDrawableProcessing(stream, n);
		r.recordDrawableCulling(stream, views.data(), V);
		std::vector<DrawRange> ranges;
		r.recordSceneRendering(stream, *ss[0], ranges);
		r.endRecording(stream);
		r.endFrame();
		if (hipStreamSynchronize(stream) != hipSuccess) throw std::runtime_error("hipStreamSynchronize");

		// the items, their spheres and the mesh table the oracle expects
		std::vector<float> spheres;
		std::vector<uint32_t> meshIndex, rangeItems;
		std::vector<std::pair<uint32_t, uint32_t>> source;
		size_t slot = 0;
		for (size_t k = 0; k < ranges.size(); k++) {
			const DrawRange& range = ranges[k];
			CHECK(range.itemBase == meshIndex.size(), "range item base");
			for (size_t j = 0; j < range.count; j++, slot++) {
				const DInfo& d = info[indexOf.at(&range.stateSet->getDrawable(j))];
				const std::vector<float>& mats = mlHost[d.ml];
				for (size_t t = 0; t < mats.size() / 16; t++) {
					DrawableGpuData rec = range.stateSet->drawableDataList()[j];
					const cadrMesh& me = expMeshes.at(rec.padding);
					const float local[4] = {me.center[0], me.center[1], me.center[2], me.radius};
					float out[4];
					oracle_sphere_transform(&mats[16 * t], local, out);
					spheres.insert(spheres.end(), out, out + 4);
					meshIndex.push_back(rec.padding);
					source.push_back({uint32_t(slot), uint32_t(t)});
				}
			}
			CHECK(range.itemCount == meshIndex.size() - range.itemBase, "range item count");
			rangeItems.push_back(uint32_t(range.itemCount));
		}
		const size_t items = meshIndex.size();
		CHECK(r.numCullItems() == items && r.numCullViews() == V, "item and view counts");
		std::vector<cadrMesh> gotMeshes(expMeshes.size());
		ds.provider().download(gotMeshes.data(), r.cullMeshTableAddress(), gotMeshes.size() * sizeof(cadrMesh));
		CHECK(std::memcmp(gotMeshes.data(), expMeshes.data(), gotMeshes.size() * sizeof(cadrMesh)) == 0,
		      "mesh table built from the PrimitiveSets");
		std::vector<float> gotSpheres(4 * items);
		ds.provider().download(gotSpheres.data(), r.cullSphereBufferAddress(), gotSpheres.size() * 4);
		CHECK(std::memcmp(gotSpheres.data(), spheres.data(), spheres.size() * 4) == 0, "item world spheres");

		std::vector<cadrDrawIndexedIndirectCommand> expCmds(V * items);
		std::vector<uint32_t> expCounts(V * ranges.size());
		oracle_cull(spheres.data(), meshIndex.data(), expMeshes.data(), uint32_t(expMeshes.size()), rangeItems.data(),
		            uint32_t(ranges.size()), views.data(), uint32_t(V), uint32_t(items), expCmds.data(),
		            expCounts.data(), 1);
		std::vector<uint32_t> gotCounts;
		r.downloadCullCounts(gotCounts);
		CHECK(gotCounts == expCounts, "per-view per-StateSet counts");
		std::vector<cadrDrawIndexedIndirectCommand> gotCmds(V * items);
		ds.provider().download(gotCmds.data(), r.cullCommandBufferAddress(), gotCmds.size() * 20);
		std::vector<cadrDrawablePointers> gather(n), gotPtr;
		ds.provider().download(gather.data(), r.drawablePointersBufferAddress(), n * 32);
		if (frame > 0) {
			gotPtr.resize(V * items);
			ds.provider().download(gotPtr.data(), r.cullPointersBufferAddress(), gotPtr.size() * 32);
		}
		size_t visible = 0;
		for (size_t w = 0; w < V; w++) {
			for (size_t k = 0; k < ranges.size(); k++) {
				for (uint32_t j = 0; j < expCounts[w * ranges.size() + k]; j++) {
					const size_t at = w * items + ranges[k].itemBase + j;
					cadrDrawIndexedIndirectCommand e = expCmds[at];
					const auto src = source.at(e.firstInstance);
					e.firstInstance = src.second;  // the matrix index of the item's MatrixList
					CHECK(std::memcmp(&gotCmds[at], &e, 20) == 0,
					      "command " + std::to_string(j) + " of view " + std::to_string(w) + " range " + std::to_string(k));
					if (frame > 0)
						CHECK(std::memcmp(&gotPtr[at], &gather[src.first], 32) == 0, "compacted DrawablePointers");
					visible++;
				}
			}
		}
		CHECK(visible > 0 && visible < V * items, "some items culled, some visible");
		if (shared) {
			// what a Vulkan renderer imports: the same bytes through the exported descriptors
			CHECK(frame == 2 || r.cullBufferGeneration() != generation, "grown buffers are a new generation");
			generation = r.cullBufferGeneration();
			auto viaImport = [&](int fd, size_t size, void* dst, size_t bytes) {
				cadrSharedBuffer* imp = nullptr;
				CHECK(fd >= 0 && cadrSharedBufferImport(fd, size, &imp) == 0, "import of an exported cull buffer");
				CHECK(hipMemcpy(dst, reinterpret_cast<void*>(cadrSharedBufferAddress(imp)), bytes,
				                hipMemcpyDeviceToHost) == hipSuccess,
				      "read through the import");
				cadrSharedBufferDestroy(imp);
			};
			std::vector<cadrDrawIndexedIndirectCommand> impCmds(V * items);
			viaImport(r.exportCullCommandBufferFd(), r.cullCommandBufferSize(), impCmds.data(), impCmds.size() * 20);
			CHECK(std::memcmp(impCmds.data(), gotCmds.data(), gotCmds.size() * 20) == 0, "imported commands");
			std::vector<uint32_t> impCounts(gotCounts.size());
			viaImport(r.exportCullCountBufferFd(), r.cullCountBufferSize(), impCounts.data(), impCounts.size() * 4);
			CHECK(impCounts == gotCounts, "imported counts");
			if (frame > 0) {
				std::vector<cadrDrawablePointers> impPtr(V * items);
				viaImport(r.exportCullPointersBufferFd(), r.cullPointersBufferSize(), impPtr.data(), impPtr.size() * 32);
				CHECK(std::memcmp(impPtr.data(), gotPtr.data(), gotPtr.size() * 32) == 0, "imported DrawablePointers");
			}
		}
	}
	(void)hipStreamDestroy(stream);
	for (void* p : deviceArrays) (void)hipFree(p);
	drawables.clear();
	mls.clear();
	geoms.clear();
	ss.clear();
}

int main(int argc, char** argv)
{
	const std::string mode = argc > 1 ? argv[1] : "cpu";
	try {
		if (mode == "cpu") {
			std::printf("handle levels\n");
			std::fflush(stdout);
			testHandleLevels();
			std::printf("allocator placement\n");
			std::fflush(stdout);
			testAllocatorPlacement();
			std::printf("allocator sequences\n");
			std::fflush(stdout);
			testAllocatorSequences();
			std::printf("allocator churn\n");
			std::fflush(stdout);
			testAllocatorChurn();
			testBufferGrowth();
			testParentChildLists();
			testStateSetOrder();
		} else if (mode == "gpu") {
			gpuEmptyFrame();
			gpuFrame(1, 500, 1);
			gpuFrame(2, 20000, 2);
			gpuFrame(3, 20000, 3);
			gpuResidentFrames(4);
			gpuResidentFrames(5);
			gpuCullFrames(6, false, 20000);
			gpuCullFrames(7, true, 8000);
			gpuCullFrames(8, true, 3000, true);
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
