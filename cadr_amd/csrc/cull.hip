// Frustum cull, LOD selection and per-StateSet stream compaction.
//
// The reference has no cull: processDrawables.comp emits every drawable and
// StateSet::recordToCommandBuffer draws StateSet k's full range
// [base_k, base_k + n_k) of the indirect buffer (src/CadR/StateSet.cpp:198-268).
// This pass keeps that range layout and writes, per view, the visible
// drawables of StateSet k compacted to the front of its range plus a draw
// count per StateSet (the operands of vkCmdDrawIndexedIndirectCount).
//
// Work is split into tiles of at most blockSize*itemsPerThread drawables that
// never cross a StateSet boundary; tiles are taken from an atomic ticket in
// array order, so a tile only ever waits on tiles that already run.  Within a
// tile, wave w owns a contiguous run of 64*itemsPerThread drawables (item j of
// lane l is drawable w*64*itemsPerThread + j*64 + l): loads are coalesced and
// the visibility of 64 drawables is one wave64 ballot.  Offsets come from a
// single-pass decoupled look-back: status words carry {epoch, state
// (aggregate / inclusive prefix), value}; the first tile of a StateSet
// publishes its inclusive prefix at once, and tiles are issued round-robin
// over StateSets, so the walks stay short.  Status words are tagged with a
// launch epoch and never cleared.
//
// Four kernels:
//   cullViewKernel - one view: one workgroup per tile, visibility kept in
//     registers, a 64-lane parallel look-back in wave 0, each visible
//     drawable's lane storing its command at its compacted slot.  It resets
//     its ticket and advances the epoch itself (the last workgroup to
//     finish), so a view costs one launch.
//   cullViewLoopKernel - the same view pass as a persistent loop over tiles,
//     the next tile's loads in flight during the current tile's look-back
//     and stores; used when a view has more tiles than resident workgroups.
//   cullFusedKernel - the same for one view, with the tile's spheres computed
//     from the last transform level instead of read (cadrPropagateCullView).
//   cullKernel - many views and many scenes (jobs) in one persistent launch:
//     a tile is read from HBM once into registers and tested against all
//     views of its job, 64 views at a time; lane v of wave 0 walks the chain
//     of view v; each visible drawable's lane stores its command directly at
//     its compacted slot.  A conservative bounding sphere of the tile, classified
//     per view with lanes as views, lets whole (tile, view) pairs skip the
//     per-drawable test when the tile is wholly outside or inside the
//     frustum.
#include "internal.h"
#include "launch.h"
#include "fp32.h"
#include "matrix_io.h"

#include <algorithm>
#include <cstddef>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace cadr {

constexpr uint32_t kViewGroup = 64;   // views whose look-backs share one wave
constexpr uint32_t kCtlTicket = 0;    // ctl words, each on its own 128-byte line
constexpr uint32_t kCtlDone = 32;
constexpr uint32_t kCtlEpoch = 64;
constexpr uint32_t kCtlError = 96;    // epoch of a launch whose look-back gave up, else 0
constexpr uint32_t kCtlWords = 128;
constexpr uint32_t kLdsMeshes = 64;
constexpr uint64_t kDefaultWaitLimitUs = 1000000;  // 1 s: some ten thousand times a tile's run time   // mesh tables up to this size are staged in LDS

constexpr uint32_t kStateAggregate = 1;
constexpr uint32_t kStatePrefix = 2;
constexpr uint32_t kValueMask = 0x3fffffffu;

// The leaf transform level of a fused propagation + one-view cull launch
// (cadrPropagateCullView): node i of the level is drawable i of the job.
struct LeafParams {
	uint64_t local;        // float mat4[numDrawables]
	uint64_t parentIndex;  // uint32[numDrawables]; unused for a root level
	uint64_t parentWorld;  // float mat4[parentCount], 0 for a root level
	uint64_t world;        // float mat4[numDrawables], written
	uint64_t spheres;      // float4[numDrawables], written if not 0
	uint32_t parentCount;
};

struct CullParams {
	const cadrCullJob* jobs;
	const uint32_t* firstWork;   // [numJobs+1] prefix of numTiles
	const uint64_t* statusBase;  // [numJobs] prefix of numViews*numTiles
	uint64_t* status;
	uint32_t* ctl;
	uint32_t numJobs;
	uint32_t totalWork;
	uint32_t splitTiles;  // batched: a job's last tiles handed out per view group
	uint32_t flags;
	uint64_t waitLimit;                 // steady-counter ticks a look-back may wait
	LeafParams leaf;                    // one-view fused kernel
	const cadrTransformLevel* leaves;   // fused batched kernel: leaf level of each job
};

// Addresses read from descriptors would be generic (FLAT) pointers, and a
// FLAT access counts on both the vector-memory and the LDS counters: every
// LDS wait would then also wait for outstanding output stores.  All device
// memory is therefore accessed through global (addrspace 1) pointers, and
// read-only descriptors (jobs, tiles, views, offsets) through constant
// (addrspace 4) pointers, which become scalar loads when the address is
// uniform.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
template <class T>
using gmut = __attribute__((address_space(1))) T*;
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ gptr<T> globalPtr(uint64_t a)
{
	return reinterpret_cast<gptr<T>>(a);
}
template <class T>
__device__ __forceinline__ gmut<T> globalMut(uint64_t a)
{
	return reinterpret_cast<gmut<T>>(a);
}
template <class T>
__device__ __forceinline__ cptr<T> constPtr(uint64_t a)
{
	return reinterpret_cast<cptr<T>>(a);
}
template <class T>
__device__ __forceinline__ uint64_t addr(const T* p)
{
	return reinterpret_cast<uint64_t>(p);
}

struct TileDesc {
	uint32_t start, count, segBase, k, chain;
	bool first, last;
};

// Loads item j = base + 64*j of a tile for every j, all issued before any is
// used: indices are clamped into the tile (a partial tile re-reads its last
// drawable), so no load sits behind a per-lane branch; an empty tile loads
// nothing.  Callers mask items at or beyond T.count.
template <int ITEMS>
__device__ __forceinline__ void loadTileItems(gptr<float4> spheres, gptr<uint32_t> meshIndex, const TileDesc& T,
                                              uint32_t base, float4 (&sp)[ITEMS], uint32_t (&mi)[ITEMS])
{
	if (T.count == 0) {
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			sp[j] = make_float4(0.f, 0.f, 0.f, 0.f);
			mi[j] = 0;
		}
		return;
	}
	const uint32_t last = T.count - 1;
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		const uint32_t e = min(base + uint32_t(j) * 64u, last);
		const gptr<float4> p = spheres + (T.start + e);
		sp[j] = make_float4(p->x, p->y, p->z, p->w);
		mi[j] = meshIndex[T.start + e];
	}
}

// Job descriptor j, field by field through a constant pointer (scalar loads).
static_assert(offsetof(cadrCullJob, numDrawables) == 56 && sizeof(cadrCullJob) == 80, "cadrCullJob layout");
__device__ __forceinline__ cadrCullJob loadJob(const cadrCullJob* jobs, uint32_t j)
{
	const uint64_t a = addr(jobs) + uint64_t(j) * sizeof(cadrCullJob);
	const cptr<uint64_t> q = constPtr<uint64_t>(a);
	const cptr<uint32_t> w = constPtr<uint32_t>(a + 56);
	cadrCullJob J;
	J.spheres = q[0]; J.meshIndex = q[1]; J.meshes = q[2]; J.tiles = q[3];
	J.views = q[4]; J.commands = q[5]; J.counts = q[6];
	J.numDrawables = w[0]; J.numTiles = w[1]; J.numStateSets = w[2];
	J.numViews = w[3]; J.commandStride = w[4]; J.numMeshes = w[5];
	return J;
}

// Leaf level of job j for the fused batched kernel (cadrTransformLevel in
// device memory, scalar loads); the parent count applies to a non-root level
// only, as in the one-view fused entry point.
static_assert(offsetof(cadrTransformLevel, numMeshes) == 64 && offsetof(cadrTransformLevel, parentCount) == 68 &&
                  sizeof(cadrTransformLevel) == 72,
              "cadrTransformLevel layout");
__device__ __forceinline__ LeafParams loadLeaf(const cadrTransformLevel* leaves, uint32_t j)
{
	const uint64_t a = addr(leaves) + uint64_t(j) * sizeof(cadrTransformLevel);
	const cptr<uint64_t> q = constPtr<uint64_t>(a);
	LeafParams L;
	L.local = q[0];
	L.parentIndex = q[1];
	L.parentWorld = q[2];
	L.world = q[3];
	L.spheres = q[7];
	L.parentCount = L.parentWorld ? constPtr<uint32_t>(a + 68)[0] : 0u;
	return L;
}

__device__ __forceinline__ TileDesc loadTile(const cadrCullJob& J, uint32_t t, uint32_t tileSize)
{
	const cptr<uint32_t> p = constPtr<uint32_t>(J.tiles) + 8 * uint64_t(t);
	const uint32_t flags = p[3];
	TileDesc d;
	d.start = p[0];
	d.count = p[1];
	d.segBase = p[2];
	d.k = flags & CADR_TILE_STATESET_MASK;
	d.first = (flags & CADR_TILE_FIRST) != 0u;
	d.last = (flags & CADR_TILE_LAST) != 0u;
	d.chain = p[4];
	if (d.chain >= J.numTiles) d.chain = t;  // malformed list: keep status accesses in range
	if (d.start >= J.numDrawables) d.count = 0;
	else if (d.count > J.numDrawables - d.start) d.count = J.numDrawables - d.start;
	if (d.count > tileSize) d.count = tileSize;
	return d;
}

// Per-view constants: scalar loads when the view index is uniform, vector
// loads when each lane takes its own view.
struct ViewC {
	float4 r0, r1, r2;
	float nearZ, farZ, sx0, sx1, sy0, sy1, lodScale, th0, th1, th2;
};

__device__ __forceinline__ ViewC loadView(uint64_t views, uint32_t v)
{
	// cadrView: viewRow[3][4], nearZ, farZ, sideX[2], sideY[2], lodScale, lodThreshold[3]
	const cptr<float> p = constPtr<float>(views + uint64_t(v) * sizeof(cadrView));
	ViewC c;
	c.r0 = make_float4(p[0], p[1], p[2], p[3]);
	c.r1 = make_float4(p[4], p[5], p[6], p[7]);
	c.r2 = make_float4(p[8], p[9], p[10], p[11]);
	c.nearZ = p[12]; c.farZ = p[13]; c.sx0 = p[14]; c.sx1 = p[15];
	c.sy0 = p[16]; c.sy1 = p[17]; c.lodScale = p[18]; c.th0 = p[19];
	c.th1 = p[20]; c.th2 = p[21];
	return c;
}

// glm mat4*vec4 order with w = 1: (m0*x + m1*y) + (m2*z + m3).
__device__ __forceinline__ float viewCoord(float4 r, float4 s)
{
	return (r.x * s.x + r.y * s.y) + (r.z * s.z + r.w);
}

// Sphere against the frustum, exactly the expression documented in cadr_hip.h.
__device__ __forceinline__ bool sphereVisible(float4 s, const ViewC& v)
{
	const float vx = viewCoord(v.r0, s);
	const float vy = viewCoord(v.r1, s);
	const float vz = viewCoord(v.r2, s);
	const bool nearOk = (vz + s.w >= v.nearZ);
	const bool farOk = (vz - s.w <= v.farZ);
	const bool xOk = (fabsf(vx) * v.sx0 - vz * v.sx1 <= s.w);
	const bool yOk = (fabsf(vy) * v.sy0 - vz * v.sy1 <= s.w);
	return nearOk & farOk & xOk & yOk;
}

// Highest LOD index of a mesh with lodCount LODs (lodCount clamped to 1..4).
__device__ __forceinline__ uint32_t maxLodOf(uint32_t lodCount)
{
	return (lodCount >= 1u ? (lodCount <= 4u ? lodCount : 4u) : 1u) - 1u;
}

__device__ __forceinline__ uint32_t lodAtDepth(float r, float vz, const ViewC& v, uint32_t maxLod)
{
	const float p = r * v.lodScale;
	const uint32_t lod = uint32_t(p < v.th0 * vz) + uint32_t(p < v.th1 * vz) + uint32_t(p < v.th2 * vz);
	return lod < maxLod ? lod : maxLod;
}

// Tile-vs-view class: 0 = test each drawable, 1 = all outside, 2 = all inside.
//
// ts is a sphere containing every drawable sphere of the tile
// (|c_i - C| + r_i <= R).  Each quantity the drawable test compares, v.z and
// g(c) = |v.x|*sideX[0] - v.z*sideX[1] (the larger of two linear plane
// functions), is Lipschitz in the sphere centre with constant L_P = the
// largest gradient norm among its linear pieces: |row2| for v.z, and
// |sideX[0]*row0 -+ sideX[1]*row2| for g (likewise for y), taken at least 1.
// A drawable's quantity then differs from the tile centre's by at most
// L*|c_i - C| <= L*(R - r_i), so the tile is wholly outside a plane when the
// centre is outside by more than L*R, and wholly inside when inside by more
// than L*R.  The norms are rounded up by a relative 2^-12; the margin m,
// about 2^-14 of the magnitudes involved, is some thousand times the FP32
// rounding of the per-drawable expressions, so the float test of every
// drawable agrees with the class.
// Bounds only: the hardware square root (within 1 ulp) is covered by the
// relative slack the callers add.
__device__ __forceinline__ float norm3(float x, float y, float z)
{
	return __builtin_amdgcn_sqrtf((x * x + y * y) + z * z);
}

__device__ __forceinline__ uint32_t classifyTile(float4 ts, const ViewC& v)
{
	const float vx = viewCoord(v.r0, ts);
	const float vy = viewCoord(v.r1, ts);
	const float vz = viewCoord(v.r2, ts);
	const float4 a = v.r0, b = v.r1, c = v.r2;
	float L = fmaxf(1.0f, norm3(c.x, c.y, c.z));
	L = fmaxf(L, norm3(v.sx0 * a.x - v.sx1 * c.x, v.sx0 * a.y - v.sx1 * c.y, v.sx0 * a.z - v.sx1 * c.z));
	L = fmaxf(L, norm3(v.sx0 * a.x + v.sx1 * c.x, v.sx0 * a.y + v.sx1 * c.y, v.sx0 * a.z + v.sx1 * c.z));
	L = fmaxf(L, norm3(v.sy0 * b.x - v.sy1 * c.x, v.sy0 * b.y - v.sy1 * c.y, v.sy0 * b.z - v.sy1 * c.z));
	L = fmaxf(L, norm3(v.sy0 * b.x + v.sy1 * c.x, v.sy0 * b.y + v.sy1 * c.y, v.sy0 * b.z + v.sy1 * c.z));
	L *= 1.0f + 1.0f / 4096.0f;
	const float R = ts.w * L * (1.0f + 1.0f / 4096.0f);
	const float mag = fabsf(ts.x) + fabsf(ts.y) + fabsf(ts.z) + ts.w + fabsf(v.r0.w) + fabsf(v.r1.w) +
	                  fabsf(v.r2.w) + fabsf(v.nearZ) + fabsf(v.farZ) + 1.0f;
	const float m = mag * L * (1.0f / 16384.0f);
	const float gx = fabsf(vx) * v.sx0 - vz * v.sx1;
	const float gy = fabsf(vy) * v.sy0 - vz * v.sy1;
	const bool out = (vz + R < v.nearZ - m) || (vz - R > v.farZ + m) || (gx > R + m) || (gy > R + m);
	if (out) return 1u;
	const bool in = (vz - R > v.nearZ + m) && (vz + R < v.farZ - m) && (gx + R < -m) && (gy + R < -m);
	return in ? 2u : 0u;
}

// Fused-multiply-add estimate of the drawable test: the smallest of the four
// inside margins (near, far, x sides, y sides; visible when all are >= 0),
// each computed with FMAs instead of the separately rounded glm expression.
// Its difference from the same margins taken on the glm-order values, plus the
// rounding that decides the glm-order comparisons, is below
// prefilterMargin(): when |planeMin| exceeds that bound its sign gives
// exactly the result of sphereVisible, else the caller evaluates
// sphereVisible.  Inputs: finite centres, radius finite >= 0 or -inf (then
// the centre is 0 and the result -inf, never visible); w0..w2 are the view's
// translation column (r0.w, r1.w, r2.w), passed separately so that a caller
// can keep them in vector registers (an FMA takes one scalar operand).
__device__ __forceinline__ float planeMin(float4 s, const ViewC& v, float w0, float w1, float w2)
{
	const float vx = __builtin_fmaf(v.r0.x, s.x, __builtin_fmaf(v.r0.y, s.y, __builtin_fmaf(v.r0.z, s.z, w0)));
	const float vy = __builtin_fmaf(v.r1.x, s.x, __builtin_fmaf(v.r1.y, s.y, __builtin_fmaf(v.r1.z, s.z, w1)));
	const float vz = __builtin_fmaf(v.r2.x, s.x, __builtin_fmaf(v.r2.y, s.y, __builtin_fmaf(v.r2.z, s.z, w2)));
	const float pn = (vz + s.w) - v.nearZ;
	const float pf = (v.farZ - vz) + s.w;
	const float px = __builtin_fmaf(-fabsf(vx), v.sx0, __builtin_fmaf(vz, v.sx1, s.w));
	const float py = __builtin_fmaf(-fabsf(vy), v.sy0, __builtin_fmaf(vz, v.sy1, s.w));
	return fminf(fminf(pn, pf), fminf(px, py));
}

// Bound for planeMin over spheres with |centre component| <= mc and radius
// <= w.  With u = 2^-24 and S_r = (|r.x|+|r.y|+|r.z|)*mc + |r.w| for view row
// r: each view coordinate differs between the FMA and the glm order by at
// most 6u*S_r; the FMA margins add at most 2u of their operands, and the glm
// comparisons are decided up to 2u of theirs, so every margin is within
// u*(10*(|side0|*S + |side1|*S_z) + 3*w + |near| + |far|) of a decisive one.
// The bound taken is 32u times the larger quantity
// ((1 + sum |side|) * (S_x+S_y+S_z) + w + |near| + |far|), plus 2^-100
// against underflow; it is NaN or infinite when a view constant is, and the
// caller uses the estimate only when the bound is at most 2^100 (no operand
// then comes near overflow).
__device__ __forceinline__ float prefilterMargin(const ViewC& v, float mc, float w)
{
	const float sx = (fabsf(v.r0.x) + fabsf(v.r0.y) + fabsf(v.r0.z)) * mc + fabsf(v.r0.w);
	const float sy = (fabsf(v.r1.x) + fabsf(v.r1.y) + fabsf(v.r1.z)) * mc + fabsf(v.r1.w);
	const float sz = (fabsf(v.r2.x) + fabsf(v.r2.y) + fabsf(v.r2.z)) * mc + fabsf(v.r2.w);
	const float sides = 1.0f + fabsf(v.sx0) + fabsf(v.sx1) + fabsf(v.sy0) + fabsf(v.sy1);
	return (sides * (sx + sy + sz) + w + fabsf(v.nearZ) + fabsf(v.farZ)) * 0x1p-19f + 0x1p-100f;
}

__device__ __forceinline__ uint32_t laneId() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

__device__ __forceinline__ uint32_t maskRank(uint64_t m)
{
	return __builtin_amdgcn_mbcnt_hi(uint32_t(m >> 32), __builtin_amdgcn_mbcnt_lo(uint32_t(m), 0u));
}

__device__ __forceinline__ uint32_t uniform(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ void waveSync()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t packStatus(uint32_t epoch, uint32_t state, uint32_t value)
{
	return (uint64_t(epoch) << 32) | (uint64_t(state) << 30) | uint64_t(value & kValueMask);
}

__device__ __forceinline__ void publish(uint64_t* p, uint64_t v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t readStatus(const uint64_t* p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool statusReady(uint64_t s, uint32_t epoch)
{
	return uint32_t(s >> 32) == epoch && (uint32_t(s) >> 30) != 0u;
}

// Look-back waits are bounded.  In a well-formed tile list every
// predecessor holds an earlier ticket and is running or done, so a wait ends
// within one tile's run time.  A caller-built list that names a predecessor
// no tile publishes (a non-FIRST tile whose chainIndex - 1 is missing) would
// wait forever; instead, a wait that lasts waitLimit ticks of the steady
// counter records the launch's epoch in the error word and gives up, and
// every later wait of the same launch gives up at its first unready poll.
// The launch then ends with unspecified outputs and the workspace reports
// CADR_ERROR_TIMEOUT (cadrCullWorkspaceCheck).  Epochs are never 0, so an
// error word of 0 means no wait gave up.
struct WaitBound {
	uint32_t* ctl;
	uint64_t limit;
	uint32_t epoch;
};

// Called after an unready poll; start is 0 until the first such poll.
__device__ __forceinline__ bool waitGiveUp(const WaitBound& B, uint64_t& start)
{
	if (__hip_atomic_load(&B.ctl[kCtlError], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == B.epoch) return true;
	const uint64_t now = uint64_t(wall_clock64());
	if (start == 0) {
		start = now == 0 ? 1 : now;
		return false;
	}
	if (now - start <= B.limit) return false;
	__hip_atomic_store(&B.ctl[kCtlError], B.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	return true;
}

// Exclusive prefix of chain position c, one chain per lane: walk back over the
// predecessors adding aggregates until an inclusive prefix is found; position
// 0 ends the walk in any case, and a wait bounded as above ends it early.
__device__ uint32_t lookBackSerial(const uint64_t* chain, uint32_t c, const WaitBound& B)
{
	uint32_t excl = 0;
	uint32_t p = c;
	uint64_t start = 0;
	while (p > 0) {
		--p;
		uint64_t s;
		for (;;) {
			s = readStatus(chain + p);
			if (statusReady(s, B.epoch)) break;
			if (waitGiveUp(B, start)) return excl;
			__builtin_amdgcn_s_sleep(1);
		}
		excl += uint32_t(s) & kValueMask;
		if ((uint32_t(s) >> 30) == kStatePrefix) break;
	}
	return excl;
}

// Exclusive prefix of chain position c with the whole wave: lane l reads
// position p-1-l, the walk stops at the nearest inclusive prefix, and 64
// predecessors cost one round trip.  Returns the same value in every lane
// (the give-up decision is made from wave-uniform values).
__device__ uint32_t lookBackWave(const uint64_t* chain, uint32_t c, const WaitBound& B, uint32_t lane)
{
	uint32_t excl = 0;
	int64_t p = c;
	uint64_t start = 0;
	while (p > 0) {
		const int64_t idx = p - 1 - int64_t(lane);
		const uint64_t s = idx >= 0 ? readStatus(chain + idx) : packStatus(B.epoch, kStatePrefix, 0);
		const bool ready = statusReady(s, B.epoch);
		const uint64_t pmask = __ballot(ready && (uint32_t(s) >> 30) == kStatePrefix);
		const uint64_t nmask = __ballot(!ready);
		const uint32_t f = pmask ? uint32_t(__builtin_ctzll(pmask)) : 64u;
		const uint64_t need = f >= 63u ? ~0ull : ((2ull << f) - 1ull);
		if (nmask & need) {
			if (uniform(uint32_t(waitGiveUp(B, start)))) return excl;
			__builtin_amdgcn_s_sleep(1);
			continue;
		}
		uint32_t v = lane <= f ? (uint32_t(s) & kValueMask) : 0u;
#pragma unroll
		for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
		excl += v;
		if (f < 64u) break;
		p -= 64;
	}
	return excl;
}

// The mesh table of a job: an LDS copy when it is small, else device memory.
// The LDS copy is field-major (dword f of mesh m at f*kLdsMeshes + m): with
// 64-byte records every mesh's copy of a field would sit in the same two LDS
// banks, and 64 lanes reading random meshes would serialise.
// cadrMesh dwords: 0-2 centre, 3 radius, 4 lodCount, 5 vertexOffset,
// 8-11 indexCount[4], 12-15 firstIndex[4]; the LDS copy holds the highest
// LOD index (maxLodOf(lodCount)) in place of dword 4.
struct MeshTable {
	const uint32_t* lds;
	gptr<cadrMesh> mem;
	bool inLds;
	__device__ __forceinline__ uint32_t field(uint32_t f, uint32_t m) const { return lds[f * kLdsMeshes + m]; }
};

// Command fields of mesh m at the LOD the view selects for a sphere of radius
// r at view depth vz.  LDS is a compile-time choice: a run-time branch
// between the two sources would make the compiler wait for every
// outstanding vector-memory operation (output stores included) where the
// branches join.
template <bool LDS>
__device__ __forceinline__ void meshCommand(const MeshTable& M, uint32_t m, float r, float vz, const ViewC& V,
                                            uint32_t& count, uint32_t& first, uint32_t& vofs)
{
	if constexpr (LDS) {
		const uint32_t lod = lodAtDepth(r, vz, V, M.field(4, m));
		count = M.field(8 + lod, m);
		first = M.field(12 + lod, m);
		vofs = M.field(5, m);
	} else {
		const gptr<cadrMesh> mesh = M.mem + m;
		const uint32_t lod = lodAtDepth(r, vz, V, maxLodOf(mesh->lodCount));
		count = mesh->indexCount[lod];
		first = mesh->firstIndex[lod];
		vofs = uint32_t(mesh->vertexOffset);
	}
}

// Local bounding sphere {centre, radius} of mesh m (m < numMeshes).
template <bool LDS>
__device__ __forceinline__ float4 meshSphere(const MeshTable& M, uint32_t m)
{
	if constexpr (LDS) {
		return make_float4(__uint_as_float(M.field(0, m)), __uint_as_float(M.field(1, m)),
		                   __uint_as_float(M.field(2, m)), __uint_as_float(M.field(3, m)));
	} else {
		return loadF4(reinterpret_cast<gptr<float4>>(M.mem + m));
	}
}

// Mesh table of a job staged in LDS, field-major (see MeshTable), when it is
// small; returns whether the LDS copy is in use.
template <int BLOCK>
__device__ __forceinline__ bool stageMeshes(uint32_t* sMeshes, const cadrCullJob& J, uint32_t tid)
{
	if (J.numMeshes > kLdsMeshes) return false;
	const gptr<uint32_t> src = globalPtr<uint32_t>(J.meshes);
	for (uint32_t i = tid; i < J.numMeshes * 16u; i += BLOCK) {
		const uint32_t x = src[i];
		sMeshes[(i % 16u) * kLdsMeshes + i / 16u] = i % 16u == 4u ? maxLodOf(x) : x;
	}
	return true;
}

// The fused form of loadTileItems: the wave's items are computed from the
// leaf level instead of read, exactly as cadrPropagateTransforms computes
// them (world = parentWorld[parentIndex[i]] * local[i], a parent index at or
// past parentCount or a root level giving world = local; the world sphere of
// mesh meshIndex[i], or the empty sphere for an index outside the table).
// World matrices (and spheres, if requested) are stored on the way.  Item j
// of the lane is tile offset base + 64*j + lane; the wave's 64 local
// matrices of item j are one contiguous 4-KiB block, loaded lane-contiguous
// (each instruction one KiB) and redistributed through LDS.  Parent and mesh
// indices of all items are loaded first (clamped into the tile, like
// loadTileItems); with PREFETCH the local block and parent matrix of item
// j+1 are issued before item j is computed, so each item's loads overlap the
// previous item's work.  The batched kernel prefetches only with 8 items per
// thread: with 4, the 32 registers of the prefetch drop it from 3 to 2 waves
// per SIMD for the whole view loop (measured slower).
struct ItemLoads {
	float4 raw[4];  // this lane's 4 float4 of the wave's 4-KiB local block
	float4 P[4];    // this lane's parent world matrix, or with COOP its 4 float4 of the
	                // wave's parent block (issueParentLoads), staged through LDS before use
};

// Every load unconditional (indices clamped), so all are in flight at once;
// lanes past the tile end load the tile's last node again and nothing of
// theirs is stored.
template <bool COOP>
__device__ __forceinline__ void issueItemLoads(const LeafParams& Lf, uint64_t first, uint64_t end, uint32_t p,
                                               uint32_t lane, ItemLoads& d)
{
	const gptr<float4> local = globalPtr<float4>(Lf.local);
#pragma unroll
	for (int k = 0; k < 4; k++) {
		const uint32_t q = uint32_t(k) * 64u + lane;  // float4 index inside the block
		const uint64_t m = min(first + q / 4u, end - 1u);
		d.raw[k] = loadF4Stream(local + (4 * m + (q & 3u)));
	}
	if constexpr (COOP) {
		if (Lf.parentCount > 0) issueParentLoads(globalPtr<float4>(Lf.parentWorld), Lf.parentCount, p, lane, d.P);
	} else {
		const uint32_t pc = p < Lf.parentCount ? p : 0u;
		if (Lf.parentCount > 0) {
			const gptr<float4> parentWorld = globalPtr<float4>(Lf.parentWorld);
#pragma unroll
			for (int c = 0; c < 4; c++) d.P[c] = loadF4(parentWorld + (4 * uint64_t(pc) + c));
		}
	}
}

// store: write the world matrices (and spheres); a tile that several work
// items of one launch share is written by one of them.
template <int ITEMS, bool LDS, bool PREFETCH = true, bool COOP = false>
__device__ __forceinline__ void propagateTileItems(const LeafParams& Lf, const cadrCullJob& J, const TileDesc& T,
                                                   uint32_t base, uint32_t lane, float4* stage,
                                                   const MeshTable& meshes, float4 (&sp)[ITEMS],
                                                   uint32_t (&mi)[ITEMS], bool store = true)
{
	const float4 empty = make_float4(0.f, 0.f, 0.f, -__builtin_inff());
	if (T.count == 0) {
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			sp[j] = empty;
			mi[j] = 0;
		}
		return;
	}
	const bool root = Lf.parentWorld == 0;
	const gptr<uint32_t> parentIndex = globalPtr<uint32_t>(Lf.parentIndex);
	const gptr<uint32_t> meshIndex = globalPtr<uint32_t>(J.meshIndex);
	const uint32_t last = T.count - 1;
	uint32_t pi[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		const uint32_t node = T.start + min(base + uint32_t(j) * 64u + lane, last);
		pi[j] = root ? ~0u : parentIndex[node];
		mi[j] = meshIndex[node];
	}
	const uint64_t end = uint64_t(T.start) + T.count;
	const gmut<float4> world = globalMut<float4>(Lf.world);
	// items whose 64 nodes all lie past the tile end come last and are empty
	const uint32_t items = base >= T.count ? 0u : min(uint32_t(ITEMS), (T.count - base + 63u) / 64u);
	ItemLoads next;
	if (PREFETCH && items > 0) issueItemLoads<COOP>(Lf, uint64_t(T.start) + base, end, pi[0], lane, next);
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		if (uint32_t(j) >= items) {
			sp[j] = empty;
			continue;
		}
		const uint64_t first = uint64_t(T.start) + base + uint32_t(j) * 64u;
		const uint64_t i = first + lane;
		const bool live = i < end;
		ItemLoads cur;
		if constexpr (PREFETCH) {
			cur = next;
			if (uint32_t(j) + 1u < items)
				issueItemLoads<COOP>(Lf, first + 64u, end, pi[j + 1 < ITEMS ? j + 1 : j], lane, next);
		} else {
			issueItemLoads<COOP>(Lf, first, end, pi[j], lane, cur);
		}
		// redistribute the local block: lane l takes matrix l
#pragma unroll
		for (int k = 0; k < 4; k++) stage[stageSlot(uint32_t(k) * 64u + lane)] = cur.raw[k];
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		float4 L[4], W[4], P[4];
#pragma unroll
		for (int c = 0; c < 4; c++) L[c] = stage[stageSlot(4 * lane + c)];
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		// the parent block the same way: lane l gets its own parent
		if constexpr (COOP) {
			if (Lf.parentCount > 0) stageMatrices(stage, lane, cur.P, P);
		} else {
#pragma unroll
			for (int c = 0; c < 4; c++) P[c] = cur.P[c];
		}
		if (live && pi[j] < Lf.parentCount) {
			mat4Mul(P, L, W);
		} else {
#pragma unroll
			for (int c = 0; c < 4; c++) W[c] = L[c];
		}
		if (store) storeMatrices(world, first, end, stage, lane, W);
		float4 s = empty;
		if (live && mi[j] < J.numMeshes) s = sphereTransform(W, meshSphere<LDS>(meshes, mi[j]));
		if (store && Lf.spheres && live) storeF4(globalMut<float4>(Lf.spheres) + i, s);
		sp[j] = s;
		// without the prefetch, keep the scheduler from hoisting later items'
		// loads above this one (every item's 32 registers of loads at once)
		if constexpr (!PREFETCH) __builtin_amdgcn_sched_barrier(0);
	}
}

// ---------------------------------------------------------------- one view

template <int BLOCK, int ITEMS, bool LDS, bool FUSED>
__device__ __forceinline__ void cullViewTile(const CullParams& P, const cadrCullJob& J, uint32_t t, uint32_t epoch,
                                             const MeshTable& meshes, uint32_t (*sCnt)[ITEMS], uint32_t* sExclP,
                                             float4* sMat, uint32_t tid, uint32_t lane, uint32_t wave)
{
	constexpr int WAVES = BLOCK / 64, WAVE_ITEMS = 64 * ITEMS;
	const TileDesc T = loadTile(J, t, BLOCK * ITEMS);
	const ViewC V = loadView(J.views, 0);
	const gptr<float4> spheres = globalPtr<float4>(J.spheres);
	const gptr<uint32_t> meshIndex = globalPtr<uint32_t>(J.meshIndex);

	// all loads first (indices clamped into the tile, so none is conditional),
	// then the tests: a wave keeps ITEMS 16-byte loads in flight at once; the
	// fused form computes the spheres from the leaf level instead
	float4 sp[ITEMS];
	uint32_t mi[ITEMS];
	if constexpr (FUSED)
		propagateTileItems<ITEMS, LDS>(P.leaf, J, T, wave * WAVE_ITEMS, lane, sMat, meshes, sp, mi);
	else
		loadTileItems<ITEMS>(spheres, meshIndex, T, wave * WAVE_ITEMS + lane, sp, mi);
	float vz[ITEMS];
	uint64_t mask[ITEMS];
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		const uint32_t e = wave * WAVE_ITEMS + j * 64 + lane;
		const bool vis = e < T.count && mi[j] < J.numMeshes && sphereVisible(sp[j], V);
		vz[j] = viewCoord(V.r2, sp[j]);
		mask[j] = __ballot(vis);
		if (lane == 0) sCnt[wave][j] = uint32_t(__builtin_popcountll(mask[j]));
	}
	__syncthreads();

	if (wave == 0) {
		uint32_t total = 0;
		for (int q = 0; q < WAVES; q++)
			for (int j = 0; j < ITEMS; j++) total += sCnt[q][j];
		uint64_t* chain = P.status;
		uint32_t excl = 0;
		if (T.first) {
			if (lane == 0) publish(chain + T.chain, packStatus(epoch, kStatePrefix, total));
		} else {
			if (lane == 0) publish(chain + T.chain, packStatus(epoch, kStateAggregate, total));
			excl = lookBackWave(chain, T.chain, WaitBound{P.ctl, P.waitLimit, epoch}, lane);
			if (lane == 0) publish(chain + T.chain, packStatus(epoch, kStatePrefix, excl + total));
		}
		if (lane == 0) {
			*sExclP = excl;
			if (T.last && T.k < J.numStateSets) globalMut<uint32_t>(J.counts)[T.k] = excl + total;
		}
	}
	__syncthreads();

	uint32_t pos = T.segBase + *sExclP;
	for (uint32_t q = 0; q < wave; q++)
		for (int j = 0; j < ITEMS; j++) pos += sCnt[q][j];
	pos = uniform(pos);
	const gmut<uint32_t> out = globalMut<uint32_t>(J.commands);
#pragma unroll
	for (int j = 0; j < ITEMS; j++) {
		const uint32_t n = uint32_t(__builtin_popcountll(mask[j]));
		if (n == 0) continue;
		if ((mask[j] >> lane) & 1ull) {
			uint32_t count, first, vofs;
			meshCommand<LDS>(meshes, mi[j], sp[j].w, vz[j], V, count, first, vofs);
			const uint32_t r = maskRank(mask[j]);
			if (pos + r < J.commandStride) {
				const gmut<uint32_t> o = out + uint64_t(pos) * 5ull + uint64_t(r * 5u);
				o[0] = count;
				o[1] = 1u;
				o[2] = first;
				o[3] = vofs;
				o[4] = T.start + wave * WAVE_ITEMS + j * 64 + lane;
			}
		}
		pos = uniform(pos + n);
	}
}


template <int BLOCK, int ITEMS, bool FUSED>
__device__ __forceinline__ void cullViewBody(const CullParams& P)
{
	constexpr int WAVES = BLOCK / 64;
	__shared__ float4 sMat[FUSED ? WAVES : 1][FUSED ? 256 : 1];
	__shared__ uint32_t sTicket;
	__shared__ uint32_t sCnt[WAVES][ITEMS];
	__shared__ uint32_t sExcl;
	__shared__ uint32_t sMeshes[16 * kLdsMeshes];

	const uint32_t tid = threadIdx.x, lane = laneId(), wave = tid / 64;
	const uint32_t epoch = __hip_atomic_load(&P.ctl[kCtlEpoch], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	if (tid == 0) sTicket = atomicAdd(&P.ctl[kCtlTicket], 1u);
	const cadrCullJob J = loadJob(P.jobs, 0);
	const bool ldsMeshes = stageMeshes<BLOCK>(sMeshes, J, tid);
	__syncthreads();
	const uint32_t t = uniform(sTicket);
	const MeshTable meshes{sMeshes, globalPtr<cadrMesh>(J.meshes), ldsMeshes};
	if (t < P.totalWork) {
		float4* mat = &sMat[FUSED ? wave : 0][0];
		if (ldsMeshes)
			cullViewTile<BLOCK, ITEMS, true, FUSED>(P, J, t, epoch, meshes, sCnt, &sExcl, mat, tid, lane, wave);
		else
			cullViewTile<BLOCK, ITEMS, false, FUSED>(P, J, t, epoch, meshes, sCnt, &sExcl, mat, tid, lane, wave);
	}

	// the last workgroup to finish resets the ticket and starts a new epoch
	__syncthreads();
	if (tid == 0) {
		const uint32_t d = atomicAdd(&P.ctl[kCtlDone], 1u);
		if (d == gridDim.x - 1) {
			atomicExch(&P.ctl[kCtlTicket], 0u);
			atomicExch(&P.ctl[kCtlDone], 0u);
			const uint32_t e = epoch + 1u;
			atomicExch(&P.ctl[kCtlEpoch], e == 0u ? 1u : e);
		}
	}
}

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void cullViewKernel(CullParams P)
{
	cullViewBody<BLOCK, ITEMS, false>(P);
}

// The one-view cull as a persistent loop: each workgroup stages the mesh
// table once and then takes tiles by ticket until none is left.  The next
// tile's ticket and its loads are issued right after the current tile's
// counts, so they are in flight during the current tile's look-back and
// stores instead of after them.  Counts and the wave offsets alternate
// between two LDS buffers, so two barriers per tile suffice.  Results are
// those of cullViewTile: the same tests, the same compaction order.
// Launched with one workgroup per resident slot when a view has more tiles
// than that (10M drawables, 1024 x 8: 45 against 52 µs); a view of fewer
// tiles takes one tile per workgroup (cullViewKernel), which saves the
// loop's extra ticket and barrier.
template <int BLOCK, int ITEMS, bool LDS>
__device__ __forceinline__ void cullViewTiles(const CullParams& P, const cadrCullJob& J, uint32_t t, uint32_t epoch,
                                              const MeshTable& meshes, uint32_t* sTicket,
                                              uint32_t (*sCnt)[BLOCK / 64][ITEMS], uint32_t* sExcl, uint32_t tid,
                                              uint32_t lane, uint32_t wave)
{
	constexpr int WAVES = BLOCK / 64, WAVE_ITEMS = 64 * ITEMS;
	const ViewC V = loadView(J.views, 0);
	const gptr<float4> spheres = globalPtr<float4>(J.spheres);
	const gptr<uint32_t> meshIndex = globalPtr<uint32_t>(J.meshIndex);
	const gmut<uint32_t> out = globalMut<uint32_t>(J.commands);
	TileDesc T = loadTile(J, t, BLOCK * ITEMS);
	float4 sp[ITEMS];
	uint32_t mi[ITEMS];
	loadTileItems<ITEMS>(spheres, meshIndex, T, wave * WAVE_ITEMS + lane, sp, mi);
	for (uint32_t b = 0;; b ^= 1u) {
		float vz[ITEMS];
		uint64_t mask[ITEMS];
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			const uint32_t e = wave * WAVE_ITEMS + j * 64 + lane;
			const bool vis = e < T.count && mi[j] < J.numMeshes && sphereVisible(sp[j], V);
			vz[j] = viewCoord(V.r2, sp[j]);
			mask[j] = __ballot(vis);
			if (lane == 0) sCnt[b][wave][j] = uint32_t(__builtin_popcountll(mask[j]));
		}
		if (tid == 0) *sTicket = atomicAdd(&P.ctl[kCtlTicket], 1u);
		__syncthreads();

		// the next tile's loads, in flight through this tile's look-back and stores
		const uint32_t tn = uniform(*sTicket);
		const bool more = tn < P.totalWork;
		TileDesc Tn{};
		float4 spN[ITEMS];
		uint32_t miN[ITEMS];
		if (more) {
			Tn = loadTile(J, tn, BLOCK * ITEMS);
			loadTileItems<ITEMS>(spheres, meshIndex, Tn, wave * WAVE_ITEMS + lane, spN, miN);
		}

		if (wave == 0) {
			uint32_t total = 0;
			for (int q = 0; q < WAVES; q++)
				for (int j = 0; j < ITEMS; j++) total += sCnt[b][q][j];
			uint64_t* chain = P.status;
			uint32_t excl = 0;
			if (T.first) {
				if (lane == 0) publish(chain + T.chain, packStatus(epoch, kStatePrefix, total));
			} else {
				if (lane == 0) publish(chain + T.chain, packStatus(epoch, kStateAggregate, total));
				excl = lookBackWave(chain, T.chain, WaitBound{P.ctl, P.waitLimit, epoch}, lane);
				if (lane == 0) publish(chain + T.chain, packStatus(epoch, kStatePrefix, excl + total));
			}
			if (lane == 0) {
				sExcl[b] = excl;
				if (T.last && T.k < J.numStateSets) globalMut<uint32_t>(J.counts)[T.k] = excl + total;
			}
		}
		__syncthreads();

		uint32_t pos = T.segBase + sExcl[b];
		for (uint32_t q = 0; q < wave; q++)
			for (int j = 0; j < ITEMS; j++) pos += sCnt[b][q][j];
		pos = uniform(pos);
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			const uint32_t n = uint32_t(__builtin_popcountll(mask[j]));
			if (n == 0) continue;
			if ((mask[j] >> lane) & 1ull) {
				uint32_t count, first, vofs;
				meshCommand<LDS>(meshes, mi[j], sp[j].w, vz[j], V, count, first, vofs);
				const uint32_t r = maskRank(mask[j]);
				if (pos + r < J.commandStride) {
					const gmut<uint32_t> o = out + uint64_t(pos) * 5ull + uint64_t(r * 5u);
					o[0] = count;
					o[1] = 1u;
					o[2] = first;
					o[3] = vofs;
					o[4] = T.start + wave * WAVE_ITEMS + j * 64 + lane;
				}
			}
			pos = uniform(pos + n);
		}
		if (!more) break;
		T = Tn;
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			sp[j] = spN[j];
			mi[j] = miN[j];
		}
	}
}

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void cullViewLoopKernel(CullParams P)
{
	constexpr int WAVES = BLOCK / 64;
	__shared__ uint32_t sTicket;
	__shared__ uint32_t sCnt[2][WAVES][ITEMS];
	__shared__ uint32_t sExcl[2];
	__shared__ uint32_t sMeshes[16 * kLdsMeshes];

	const uint32_t tid = threadIdx.x, lane = laneId(), wave = tid / 64;
	const uint32_t epoch = __hip_atomic_load(&P.ctl[kCtlEpoch], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	if (tid == 0) sTicket = atomicAdd(&P.ctl[kCtlTicket], 1u);
	const cadrCullJob J = loadJob(P.jobs, 0);
	const bool ldsMeshes = stageMeshes<BLOCK>(sMeshes, J, tid);
	__syncthreads();
	const uint32_t t = uniform(sTicket);
	__syncthreads();  // every wave has read the first ticket before the loop rewrites it
	const MeshTable meshes{sMeshes, globalPtr<cadrMesh>(J.meshes), ldsMeshes};
	if (t < P.totalWork) {
		if (ldsMeshes)
			cullViewTiles<BLOCK, ITEMS, true>(P, J, t, epoch, meshes, &sTicket, sCnt, sExcl, tid, lane, wave);
		else
			cullViewTiles<BLOCK, ITEMS, false>(P, J, t, epoch, meshes, &sTicket, sCnt, sExcl, tid, lane, wave);
	}

	// the last workgroup to finish resets the ticket and starts a new epoch
	__syncthreads();
	if (tid == 0) {
		const uint32_t d = atomicAdd(&P.ctl[kCtlDone], 1u);
		if (d == gridDim.x - 1) {
			atomicExch(&P.ctl[kCtlTicket], 0u);
			atomicExch(&P.ctl[kCtlDone], 0u);
			const uint32_t e = epoch + 1u;
			atomicExch(&P.ctl[kCtlEpoch], e == 0u ? 1u : e);
		}
	}
}

// The fused kernel is latency-bound (a chain of dependent loads per item), so
// occupancy pays: WPE waves per SIMD is requested from the register allocator
// (4 for the small item counts, where it costs no spills).  Larger workgroups
// give each wave fewer items, so the chain per tile is shorter.
template <int BLOCK, int ITEMS, int WPE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(WPE))) void cullFusedKernel(CullParams P)
{
	cullViewBody<BLOCK, ITEMS, true>(P);
}

// ------------------------------------------------------- many views, jobs

template <int BLOCK, int ITEMS>
struct CullShared {
	static constexpr int WAVES = BLOCK / 64;
	static constexpr int TILE = BLOCK * ITEMS;

	uint32_t cnt[kViewGroup][WAVES];
	uint32_t base[kViewGroup][WAVES];  // a wave's first output slot in its view's StateSet segment

	uint32_t meshes[16 * kLdsMeshes];
	float red[WAVES][6];
	float red2[WAVES];
	uint32_t bad[WAVES];
	uint32_t ticket;
};

// COOP (fused form only): parent matrices read whole, 16 per instruction
// (issueParentLoads), for leaves whose parents are scattered; it holds more
// registers, so it is chosen per shape (see the kernel table).
template <int BLOCK, int ITEMS, bool FUSED, bool COOP = false>
__global__ __launch_bounds__(BLOCK) void cullKernel(CullParams P)
{
	using S = CullShared<BLOCK, ITEMS>;
	constexpr int WAVES = S::WAVES;
	constexpr int WAVE_ITEMS = 64 * ITEMS;
	__shared__ S sh;
	__shared__ float4 sMat[FUSED ? WAVES : 1][FUSED ? 256 : 1];  // local-matrix staging of the fused form

	const uint32_t tid = threadIdx.x;
	const uint32_t lane = laneId();
	const uint32_t wave = tid / 64;
	const uint32_t epoch = __hip_atomic_load(&P.ctl[kCtlEpoch], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
	uint32_t job = 0, stagedJob = ~0u;

	for (;;) {
		if (tid == 0) sh.ticket = atomicAdd(&P.ctl[kCtlTicket], 1u);
		__syncthreads();
		const uint32_t w = uniform(sh.ticket);
		if (w >= P.totalWork) break;

		// job of this ticket: tickets only grow, so search upward from the last job
		{
			uint32_t lo = job, hi = P.numJobs - 1;
			while (lo < hi) {
				const uint32_t mid = (lo + hi + 1) / 2;
				if (constPtr<uint32_t>(addr(P.firstWork))[mid] <= w) lo = mid; else hi = mid - 1;
			}
			job = uniform(lo);
		}
		const cadrCullJob J = loadJob(P.jobs, job);
		// the work item: a whole tile for every view, or, among the job's last
		// split tiles, one tile for one group of views (tile-major, so a
		// tile's look-back predecessors still hold earlier tickets)
		const uint32_t u = w - constPtr<uint32_t>(addr(P.firstWork))[job];
		const uint32_t groups = J.numViews / kViewGroup + (J.numViews % kViewGroup != 0u ? 1u : 0u);
		const uint32_t split = groups > 1 ? min(J.numTiles, P.splitTiles) : 0u;
		const uint32_t whole = J.numTiles - split;
		uint32_t t = u, gBegin = 0, gEnd = J.numViews;
		if (u >= whole) {
			const uint32_t q = u - whole;
			t = uniform(whole + q / groups);
			gBegin = uniform((q % groups) * kViewGroup);
			gEnd = gBegin + min(J.numViews - gBegin, kViewGroup);
		}
		const bool storeLeaf = gBegin == 0;
		const TileDesc T = loadTile(J, t, S::TILE);
		// the job's mesh table is staged once per run of tickets of one job
		// (every wave left the previous tile's last barrier, so no wave still
		// reads the old table)
		bool ldsMeshes = J.numMeshes <= kLdsMeshes;
		if (job != stagedJob) {
			ldsMeshes = stageMeshes<BLOCK>(sh.meshes, J, tid);
			stagedJob = job;
		}

		const MeshTable meshes{sh.meshes, globalPtr<cadrMesh>(J.meshes), ldsMeshes};

		// load the tile into registers (wave-contiguous runs) and bound it for
		// the tile sphere; the fused form computes the spheres from the job's
		// leaf level instead, writing its world matrices on the way
		float4 sp[ITEMS];
		uint32_t mi[ITEMS];
		if constexpr (FUSED) {
			__syncthreads();  // the sphere transforms read the LDS mesh table
			const LeafParams Lf = loadLeaf(P.leaves, job);
			constexpr bool PREFETCH = ITEMS >= 8;
			float4* stage = &sMat[wave][0];
			if (ldsMeshes)
				propagateTileItems<ITEMS, true, PREFETCH, COOP>(Lf, J, T, wave * WAVE_ITEMS, lane, stage, meshes, sp,
				                                                mi, storeLeaf);
			else
				propagateTileItems<ITEMS, false, PREFETCH, COOP>(Lf, J, T, wave * WAVE_ITEMS, lane, stage, meshes, sp,
				                                                 mi, storeLeaf);
		} else {
			loadTileItems<ITEMS>(globalPtr<float4>(J.spheres), globalPtr<uint32_t>(J.meshIndex), T,
			                     wave * WAVE_ITEMS + lane, sp, mi);
		}
		bool irregular = false;
		float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
		float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			const uint32_t e = wave * WAVE_ITEMS + j * 64 + lane;
			float4 s = sp[j];
			uint32_t m = mi[j];
			if (e >= T.count) {
				s = make_float4(0.f, 0.f, 0.f, -__builtin_inff());
				m = 0;
			} else if (m >= J.numMeshes) {
				s = make_float4(0.f, 0.f, 0.f, -__builtin_inff());  // invalid mesh: never visible
				m = 0;
			}
			sp[j] = s;
			mi[j] = m;

			// a sphere other than a finite one with r >= 0 or the empty sphere
			// (r = -inf, never visible) disables the tile sphere for this tile
			if (e < T.count && s.w != -__builtin_inff() &&
			    !(__builtin_isfinite(s.x) && __builtin_isfinite(s.y) && __builtin_isfinite(s.z) &&
			      __builtin_isfinite(s.w) && s.w >= 0.f))
				irregular = true;
			if (s.w >= 0.f) {
				mn[0] = fminf(mn[0], s.x - s.w); mx[0] = fmaxf(mx[0], s.x + s.w);
				mn[1] = fminf(mn[1], s.y - s.w); mx[1] = fmaxf(mx[1], s.y + s.w);
				mn[2] = fminf(mn[2], s.z - s.w); mx[2] = fmaxf(mx[2], s.z + s.w);
			}
		}
		const bool useTileSphere = (P.flags & CADR_CULL_NO_TILE_SPHERES) == 0u;
		if (useTileSphere) {
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
				for (int a = 0; a < 3; a++) {
					mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
					mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
				}
			}
			const bool waveIrregular = __ballot(irregular) != 0ull;
			if (lane == 0) {
				sh.bad[wave] = waveIrregular ? 1u : 0u;
				sh.red[wave][0] = mn[0]; sh.red[wave][1] = mn[1]; sh.red[wave][2] = mn[2];
				sh.red[wave][3] = mx[0]; sh.red[wave][4] = mx[1]; sh.red[wave][5] = mx[2];
			}
		}
		__syncthreads();
		bool tileUsable = false;
		float4 tileSphere = make_float4(0.f, 0.f, 0.f, -1.f);
		float tileCoord = 0.f;  // bound on |centre component| over the tile
		if (useTileSphere) {
			uint32_t anyBad = 0;
#pragma unroll
			for (int q = 0; q < WAVES; q++) {
				anyBad |= sh.bad[q];
				for (int a = 0; a < 3; a++) {
					mn[a] = fminf(mn[a], sh.red[q][a]);
					mx[a] = fmaxf(mx[a], sh.red[q][3 + a]);
				}
			}
			// centre: the middle of the tile's bounding box; radius: the largest
			// |c_i - C| + r_i over the tile (a second reduction), which for
			// scattered drawables is up to sqrt(3) times tighter than the box's
			// half diagonal
			const float cx = (mn[0] + mx[0]) * 0.5f, cy = (mn[1] + mx[1]) * 0.5f, cz = (mn[2] + mx[2]) * 0.5f;
			float rad = 0.f;
#pragma unroll
			for (int j = 0; j < ITEMS; j++) {
				const float4 q = sp[j];
				if (q.w >= 0.f) rad = fmaxf(rad, norm3(q.x - cx, q.y - cy, q.z - cz) + q.w);
			}
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1) rad = fmaxf(rad, __shfl_xor(rad, off));
			if (lane == 0) sh.red2[wave] = rad;
			__syncthreads();
			float R = 0.f;
#pragma unroll
			for (int q = 0; q < WAVES; q++) R = fmaxf(R, sh.red2[q]);
			R *= 1.0f + 1.0f / 4096.0f;
			tileUsable = anyBad == 0u && __builtin_isfinite(R) && __builtin_isfinite(cx) &&
			             __builtin_isfinite(cy) && __builtin_isfinite(cz);
			// inflate against rounding of the centre and radius above
			const float slack = (fabsf(cx) + fabsf(cy) + fabsf(cz) + R) * (1.0f / 65536.0f);
			tileSphere = make_float4(cx, cy, cz, R + slack);
			tileCoord = fmaxf(fmaxf(fmaxf(fabsf(mn[0]), fabsf(mx[0])), fmaxf(fabsf(mn[1]), fabsf(mx[1]))),
			                  fmaxf(fabsf(mn[2]), fabsf(mx[2])));
		}
		const bool emptyTile = T.count == 0;

		const uint64_t views = J.views;
		uint64_t* chain0 = P.status + constPtr<uint64_t>(addr(P.statusBase))[job];

		for (uint32_t g0 = gBegin; g0 < gEnd; g0 += kViewGroup) {
			const uint32_t vg = uniform(min(kViewGroup, J.numViews - g0));

			// tile classes of the group's views, lane v holding view g0+v, and
			// the views that see part of the tile (every wave computes the same
			// classes, so the mask is the same in every wave); the others
			// contribute nothing and are skipped by all three phases
			// (and the bound of the FMA estimate for the view; NaN: not used)
			uint32_t cls = 0;
			float margin = __builtin_nanf("");
			if (emptyTile) {
				cls = 1;
			} else if (tileUsable && lane < vg) {
				const ViewC Vl = loadView(views, g0 + lane);
				cls = classifyTile(tileSphere, Vl);
				margin = prefilterMargin(Vl, tileCoord, tileSphere.w);
			}
			const uint64_t active = __ballot(lane < vg && cls != 1u);

			// 1. visibility bits and per-wave counts.  Lanes past the tile and
			// invalid meshes hold the empty sphere, which no test passes.  Each
			// lane keeps its own bits: bit v%32 of vlo[j] (views 0-31 of the
			// group) or vhi[j] (views 32-63) is item j's visibility in view v.
			uint32_t vlo[ITEMS], vhi[ITEMS];
#pragma unroll
			for (int j = 0; j < ITEMS; j++) vlo[j] = vhi[j] = 0u;
			for (uint64_t am = active; am != 0ull; am &= am - 1ull) {
				const uint32_t v = uint32_t(__builtin_ctzll(am));
				const uint32_t c = uniform(__builtin_amdgcn_readlane(cls, v));
				const float mg = __uint_as_float(uniform(__builtin_amdgcn_readlane(__float_as_uint(margin), v)));
				const ViewC V = loadView(views, g0 + v);
				bool vis[ITEMS];
				if (c == 2u) {
#pragma unroll
					for (int j = 0; j < ITEMS; j++) vis[j] = sp[j].w >= 0.f;
				} else if (mg <= 0x1p100f) {
					// the translation column in vector registers once per view
					float w0 = V.r0.w, w1 = V.r1.w, w2 = V.r2.w;
					asm volatile("" : "+v"(w0), "+v"(w1), "+v"(w2));
#pragma unroll
					for (int j = 0; j < ITEMS; j++) {
						const float p = planeMin(sp[j], V, w0, w1, w2);
						vis[j] = p > 0.f;
						const bool unsure = !(fabsf(p) > mg);
						if (__ballot(unsure) != 0ull) vis[j] = unsure ? sphereVisible(sp[j], V) : vis[j];
					}
				} else {
#pragma unroll
					for (int j = 0; j < ITEMS; j++) vis[j] = sphereVisible(sp[j], V);
				}
				uint32_t n = 0;
				const uint32_t bit = 1u << (v & 31u);
				if (v < 32u) {
#pragma unroll
					for (int j = 0; j < ITEMS; j++) vlo[j] |= vis[j] ? bit : 0u;
				} else {
#pragma unroll
					for (int j = 0; j < ITEMS; j++) vhi[j] |= vis[j] ? bit : 0u;
				}
#pragma unroll
				for (int j = 0; j < ITEMS; j++) n += uint32_t(__builtin_popcountll(__ballot(vis[j])));
				if (lane == 0) sh.cnt[v][wave] = n;
			}
			__syncthreads();

			// 2. look-back, one view per lane of wave 0
			if (wave == 0 && lane < vg) {
				const bool act = (active >> lane) & 1ull;
				uint32_t total = 0;
				if (act) {
#pragma unroll
					for (int q = 0; q < WAVES; q++) total += sh.cnt[lane][q];
				}
				uint64_t* chain = chain0 + uint64_t(g0 + lane) * J.numTiles;
				uint32_t excl = 0;
				if (T.first) {
					publish(chain + T.chain, packStatus(epoch, kStatePrefix, total));
				} else {
					publish(chain + T.chain, packStatus(epoch, kStateAggregate, total));
					excl = lookBackSerial(chain, T.chain, WaitBound{P.ctl, P.waitLimit, epoch});
					publish(chain + T.chain, packStatus(epoch, kStatePrefix, excl + total));
				}
				if (act) {
					uint32_t b = T.segBase + excl;
#pragma unroll
					for (int q = 0; q < WAVES; q++) {
						sh.base[lane][q] = b;
						b += sh.cnt[lane][q];
					}
				}
				if (T.last && T.k < J.numStateSets)
					globalMut<uint32_t>(J.counts)[uint64_t(g0 + lane) * J.numStateSets + T.k] = excl + total;
			}
			__syncthreads();

			// 3. compacted writes: each visible drawable's lane stores its command
			// at the wave's output slot plus its rank among the item's visible
			// lanes; consecutive visible lanes write consecutive commands, so a
			// store instruction covers one contiguous run and the L2 merges the
			// runs of an item into whole lines.  Views 0-31 and 32-63 of the
			// group take their bits from vlo and vhi.
			auto writeHalf = [&](auto lds, uint64_t am, uint32_t v0,
			                     const uint32_t (&bits)[ITEMS]) __attribute__((always_inline)) {
				for (; am != 0ull; am &= am - 1ull) {
					const uint32_t v = v0 + uint32_t(__builtin_ctzll(am));
					const uint32_t run = uniform(sh.cnt[v][wave]);
					if (run == 0) continue;
					const ViewC V = loadView(views, g0 + v);
					const gmut<uint32_t> out = globalMut<uint32_t>(J.commands) + uint64_t(g0 + v) * J.commandStride * 5ull;
					const uint32_t base = uniform(sh.base[v][wave]);
					// slot bounds per lane only when the wave's run can reach the stride
					auto items = [&](auto checked) __attribute__((always_inline)) {
						uint32_t pos = base;
#pragma unroll
						for (int j = 0; j < ITEMS; j++) {
							const bool vis = (bits[j] >> (v & 31u)) & 1u;
							const uint64_t m = __ballot(vis);
							const uint32_t n = uint32_t(__builtin_popcountll(m));
							if (n == 0) continue;
							if (vis) {
								uint32_t count, first, vofs;
								meshCommand<decltype(lds)::value>(meshes, mi[j], sp[j].w, viewCoord(V.r2, sp[j]), V,
								                                  count, first, vofs);
								const uint32_t r = maskRank(m);
								if (!decltype(checked)::value || pos + r < J.commandStride) {
									const gmut<uint32_t> o = out + uint64_t(pos) * 5ull + uint64_t(r * 5u);
									o[0] = count;
									o[1] = 1u;
									o[2] = first;
									o[3] = vofs;
									o[4] = T.start + wave * WAVE_ITEMS + j * 64 + lane;
								}
							}
							pos = uniform(pos + n);
						}
					};
					if (base + run <= J.commandStride)
						items(std::false_type{});
					else
						items(std::true_type{});
				}
			};
			auto writeViews = [&](auto lds) __attribute__((always_inline)) {
				writeHalf(lds, active & 0xffffffffull, 0u, vlo);
				writeHalf(lds, active >> 32, 32u, vhi);
			};
			if (meshes.inLds)
				writeViews(std::true_type{});
			else
				writeViews(std::false_type{});
			__syncthreads();
		}
	}

	// the last workgroup to leave resets the ticket and retires this launch's
	// epoch, so the next launch on the workspace (of either kernel) never sees
	// this launch's status words as its own
	if (tid == 0) {
		const uint32_t d = atomicAdd(&P.ctl[kCtlDone], 1u);
		if (d == gridDim.x - 1) {
			atomicExch(&P.ctl[kCtlTicket], 0u);
			atomicExch(&P.ctl[kCtlDone], 0u);
			const uint32_t e = epoch + 1u;
			atomicExch(&P.ctl[kCtlEpoch], e == 0u ? 1u : e);
		}
	}
}

// Work items of a job in a batched launch: its tiles, plus one more per
// extra view group for each of its last splitTiles tiles (the kernel's rule).
// The host checks that a launch's total fits 32 bits.
__host__ __device__ __forceinline__ uint64_t jobWorkItems(uint32_t numTiles, uint32_t numViews, uint32_t splitTiles)
{
	const uint64_t groups = (uint64_t(numViews) + kViewGroup - 1) / kViewGroup;
	const uint64_t split = groups > 1 ? (numTiles < splitTiles ? numTiles : splitTiles) : 0u;
	return numTiles + split * (groups - 1);
}

// Job offsets for a batched launch (prefix sums over jobs), ticket reset and
// a new epoch; one block, before the cull kernel on the same stream.
__global__ __launch_bounds__(256) void cullSetupKernel(const cadrCullJob* __restrict__ jobs, uint32_t numJobs,
                                                       uint32_t splitTiles, uint32_t* __restrict__ firstWork,
                                                       uint64_t* __restrict__ statusBase, uint32_t* __restrict__ ctl)
{
	__shared__ uint32_t sT[256];
	__shared__ uint64_t sS[256];
	const uint32_t tid = threadIdx.x;
	uint32_t runT = 0;
	uint64_t runS = 0;
	for (uint32_t base = 0; base < numJobs; base += 256) {
		const uint32_t j = base + tid;
		uint32_t nt = 0;
		uint64_t ns = 0;
		if (j < numJobs) {
			nt = uint32_t(jobWorkItems(jobs[j].numTiles, jobs[j].numViews, splitTiles));
			ns = uint64_t(jobs[j].numViews) * jobs[j].numTiles;
		}
		sT[tid] = nt;
		sS[tid] = ns;
		__syncthreads();
		for (uint32_t off = 1; off < 256; off <<= 1) {
			const uint32_t a = tid >= off ? sT[tid - off] : 0u;
			const uint64_t b = tid >= off ? sS[tid - off] : 0ull;
			__syncthreads();
			sT[tid] += a;
			sS[tid] += b;
			__syncthreads();
		}
		if (j < numJobs) {
			firstWork[j] = runT + sT[tid] - nt;
			statusBase[j] = runS + sS[tid] - ns;
		}
		runT += sT[255];
		runS += sS[255];
		__syncthreads();
	}
	if (tid == 0) {
		firstWork[numJobs] = runT;
		ctl[kCtlTicket] = 0;
		ctl[kCtlDone] = 0;
		uint32_t e = ctl[kCtlEpoch] + 1u;
		ctl[kCtlEpoch] = e == 0u ? 1u : e;
	}
}

}  // namespace cadr

struct cadrCullWorkspace {
	uint64_t* status = nullptr;
	uint64_t statusWords = 0;
	uint32_t* ctl = nullptr;
	uint32_t* firstWork = nullptr;
	uint64_t* statusBase = nullptr;
	uint32_t maxJobs = 0;
	int device = 0;
	uint64_t waitLimit = 0;      // steady-counter ticks a look-back wait may last
	uint64_t ticksPerUs = 100;   // steady-counter rate
};

using namespace cadr;

namespace {

using KernelFn = void (*)(CullParams);

struct KernelEntry {
	uint32_t block, items;
	KernelFn batch;       // many views / jobs
	KernelFn view;        // one view, or nullptr
	KernelFn fused;       // leaf propagation + one view, or nullptr
	KernelFn fusedBatch;  // leaf propagation + many views / jobs, or nullptr
	KernelFn viewLoop;    // one view, persistent (cullViewLoopKernel), or nullptr where it is not faster
};

// viewLoop where it measured faster at 10M drawables (one view of the full10m
// scene): not at 512 x 2/8/16, and 1024 x 16's loop form spills registers.
// fusedBatch reads parents whole (COOP) at 256 x 8 and 512 x 4, where it
// measured faster on uneven64 (2.39 -> 2.32 and 2.93 -> 2.84 ms); at 256 x 4
// its registers cost a wave per SIMD (2.38 -> 2.60 ms)
const KernelEntry kKernels[] = {
    {256, 2, cullKernel<256, 2, false>, cullViewKernel<256, 2>, cullFusedKernel<256, 2, 4>, nullptr, cullViewLoopKernel<256, 2>},
    {256, 4, cullKernel<256, 4, false>, cullViewKernel<256, 4>, cullFusedKernel<256, 4, 4>, cullKernel<256, 4, true>, cullViewLoopKernel<256, 4>},
    {256, 8, cullKernel<256, 8, false>, cullViewKernel<256, 8>, cullFusedKernel<256, 8, 1>, cullKernel<256, 8, true, true>, cullViewLoopKernel<256, 8>},
    {256, 16, cullKernel<256, 16, false>, cullViewKernel<256, 16>, cullFusedKernel<256, 16, 1>, nullptr, cullViewLoopKernel<256, 16>},
    {512, 2, cullKernel<512, 2, false>, cullViewKernel<512, 2>, cullFusedKernel<512, 2, 1>, nullptr, nullptr},
    {512, 4, cullKernel<512, 4, false>, cullViewKernel<512, 4>, cullFusedKernel<512, 4, 1>, cullKernel<512, 4, true, true>, cullViewLoopKernel<512, 4>},
    {512, 8, nullptr, cullViewKernel<512, 8>, nullptr, nullptr, nullptr},
    {512, 16, nullptr, cullViewKernel<512, 16>, nullptr, nullptr, nullptr},
    {1024, 1, nullptr, cullViewKernel<1024, 1>, cullFusedKernel<1024, 1, 1>, nullptr, cullViewLoopKernel<1024, 1>},
    {1024, 2, nullptr, cullViewKernel<1024, 2>, cullFusedKernel<1024, 2, 1>, nullptr, cullViewLoopKernel<1024, 2>},
    {1024, 4, nullptr, cullViewKernel<1024, 4>, nullptr, nullptr, cullViewLoopKernel<1024, 4>},
    {1024, 8, nullptr, cullViewKernel<1024, 8>, nullptr, nullptr, cullViewLoopKernel<1024, 8>},
    {1024, 16, nullptr, cullViewKernel<1024, 16>, nullptr, nullptr, nullptr},
    {128, 8, cullKernel<128, 8, false>, nullptr, nullptr, nullptr, nullptr},
    {128, 16, cullKernel<128, 16, false>, nullptr, nullptr, nullptr, nullptr},
};

// Workgroups of kernel fn (block threads) resident on the current device at
// once: compute units x workgroups per CU by the occupancy calculator,
// cached per (device, kernel).  0 if the query fails.
int residentWorkgroups(KernelFn fn, uint32_t block)
{
	static std::mutex m;
	static std::vector<std::pair<std::pair<int, KernelFn>, int>> cache;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return 0;
	std::lock_guard<std::mutex> lock(m);
	for (const auto& e : cache)
		if (e.first.first == dev && e.first.second == fn) return e.second;
	int perCU = 0;
	if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, reinterpret_cast<const void*>(fn), int(block), 0) !=
	        hipSuccess ||
	    perCU <= 0)
		return 0;
	const int n = computeUnitCount() * perCU;
	if (n > 0) cache.push_back({{dev, fn}, n});
	return n;
}

const KernelEntry* findKernel(const cadrCullConfig* c)
{
	if (!c) return nullptr;
	for (const KernelEntry& k : kKernels)
		if (k.block == c->blockSize && k.items == c->itemsPerThread) return &k;
	return nullptr;
}

}  // namespace

bool cadr::cullHasKernel(const cadrCullConfig* c, CullPath path)
{
	const KernelEntry* k = findKernel(c);
	if (!k) return false;
	switch (path) {
	case CullPath::View: return k->view != nullptr;
	case CullPath::Batch: return k->batch != nullptr;
	case CullPath::Fused: return k->fused != nullptr;
	case CullPath::FusedBatch: return k->fusedBatch != nullptr;
	}
	return false;
}

namespace {

int checkJob(const cadrCullJob& J)
{
	CADR_REQUIRE(J.numTiles == 0 || J.tiles, "cull: job without tiles");
	CADR_REQUIRE(J.numTiles == 0 || (J.spheres && J.meshIndex && J.meshes && J.views && J.commands && J.counts),
	             "cull: job with a null array");
	CADR_REQUIRE(J.spheres % 16 == 0 && J.meshes % 16 == 0 && J.views % 16 == 0 && J.tiles % 16 == 0 &&
	                 J.meshIndex % 4 == 0 && J.commands % 4 == 0 && J.counts % 4 == 0,
	             "cull: misaligned job array (spheres, meshes, views, tiles need 16 bytes)");
	CADR_REQUIRE(J.commandStride >= J.numDrawables, "cull: commandStride < numDrawables");
	return CADR_SUCCESS;
}

}  // namespace

CADR_API uint32_t cadrCullTileSize(const cadrCullConfig* config)
{
	const KernelEntry* k = findKernel(config);
	return k ? k->block * k->items : 0u;
}

CADR_API uint32_t cadrCullConfigPaths(const cadrCullConfig* config)
{
	const KernelEntry* k = findKernel(config);
	if (!k) return 0u;
	return (k->view || k->batch ? CADR_CULL_PATH_VIEW : 0u) | (k->batch ? CADR_CULL_PATH_BATCH : 0u) |
	       (k->fused ? CADR_CULL_PATH_FUSED_VIEW : 0u) | (k->fusedBatch ? CADR_CULL_PATH_FUSED_BATCH : 0u);
}

CADR_API int cadrBuildCullTiles(const uint32_t* stateSetCounts, uint32_t numStateSets, uint32_t tileSize,
                                cadrCullTile* tiles, uint32_t capacity, uint32_t* numTiles)
{
	CADR_REQUIRE(numTiles, "cadrBuildCullTiles: numTiles is NULL");
	CADR_REQUIRE(tileSize > 0, "cadrBuildCullTiles: tileSize is 0");
	CADR_REQUIRE(numStateSets == 0 || stateSetCounts, "cadrBuildCullTiles: stateSetCounts is NULL");
	CADR_REQUIRE(numStateSets <= CADR_TILE_STATESET_MASK, "cadrBuildCullTiles: too many StateSets");
	std::vector<uint64_t> base(numStateSets + 1, 0), firstTile(numStateSets + 1, 0);
	uint64_t maxTiles = 0;
	for (uint32_t k = 0; k < numStateSets; k++) {
		const uint64_t c = stateSetCounts[k];
		// look-back status words carry 30-bit counts
		if (c > CADR_TILE_STATESET_MASK)
			return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: a StateSet holds more than 2^30-1 drawables");
		const uint64_t nt = c == 0 ? 1 : (c + tileSize - 1) / tileSize;
		base[k + 1] = base[k] + c;
		firstTile[k + 1] = firstTile[k] + nt;
		maxTiles = std::max(maxTiles, nt);
	}
	if (base[numStateSets] > 0xffffffffull)
		return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: more than 2^32 drawables");
	const uint64_t n = firstTile[numStateSets];
	if (n > 0xffffffffull) return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: too many tiles");
	*numTiles = uint32_t(n);
	if (!tiles) return ok();
	if (n > capacity) return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: capacity too small");
	// round-robin over StateSets: the q-th tile of every StateSet, then the (q+1)-th
	uint64_t out = 0;
	for (uint64_t q = 0; q < maxTiles; q++) {
		for (uint32_t k = 0; k < numStateSets; k++) {
			const uint64_t c = stateSetCounts[k];
			const uint64_t nt = firstTile[k + 1] - firstTile[k];
			if (q >= nt) continue;
			const uint64_t cnt = c == 0 ? 0 : std::min<uint64_t>(tileSize, c - q * tileSize);
			uint32_t flags = k;
			if (q == 0) flags |= CADR_TILE_FIRST;
			if (q + 1 == nt) flags |= CADR_TILE_LAST;
			cadrCullTile& t = tiles[out++];
			t.start = uint32_t(base[k] + q * tileSize);
			t.count = uint32_t(cnt);
			t.stateSetBase = uint32_t(base[k]);
			t.stateSetAndFlags = flags;
			t.chainIndex = uint32_t(firstTile[k] + q);
			t.reserved[0] = t.reserved[1] = t.reserved[2] = 0;
		}
	}
	return ok();
}

CADR_API int cadrCullWorkspaceCreate(uint64_t statusWords, uint32_t maxJobs, cadrCullWorkspace** workspace)
{
	CADR_REQUIRE(workspace, "cadrCullWorkspaceCreate: workspace is NULL");
	CADR_REQUIRE(maxJobs > 0, "cadrCullWorkspaceCreate: maxJobs is 0");
	*workspace = nullptr;
	auto* w = new cadrCullWorkspace;
	CADR_CHECK_HIP(hipGetDevice(&w->device), "hipGetDevice");
	w->statusWords = statusWords > 0 ? statusWords : 1;
	w->maxJobs = maxJobs;
	hipError_t e = hipMalloc(&w->status, w->statusWords * sizeof(uint64_t));
	if (e == hipSuccess) e = hipMalloc(&w->ctl, kCtlWords * sizeof(uint32_t));
	if (e == hipSuccess) e = hipMalloc(&w->firstWork, (uint64_t(maxJobs) + 1) * sizeof(uint32_t));
	if (e == hipSuccess) e = hipMalloc(&w->statusBase, uint64_t(maxJobs) * sizeof(uint64_t));
	// epoch 0 is never a valid epoch, so zeroed status words read as "not ready"
	if (e == hipSuccess) e = hipMemset(w->status, 0, w->statusWords * sizeof(uint64_t));
	if (e == hipSuccess) e = hipMemset(w->ctl, 0, kCtlWords * sizeof(uint32_t));
	// launches start at epoch 1: 0 marks a status word never written and an
	// error word with nothing to report
	const uint32_t firstEpoch = 1;
	if (e == hipSuccess) e = hipMemcpy(w->ctl + kCtlEpoch, &firstEpoch, sizeof(firstEpoch), hipMemcpyHostToDevice);
	int khz = 0;
	if (e == hipSuccess) e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, w->device);
	w->ticksPerUs = khz >= 1000 ? uint64_t(khz) / 1000u : 1u;
	w->waitLimit = kDefaultWaitLimitUs * w->ticksPerUs;
	if (e == hipSuccess) e = hipDeviceSynchronize();
	if (e != hipSuccess) {
		(void)hipGetLastError();
		(void)hipFree(w->status); (void)hipFree(w->ctl); (void)hipFree(w->firstWork); (void)hipFree(w->statusBase);
		delete w;
		return setError(CADR_ERROR_OUT_OF_RESOURCES, "cadrCullWorkspaceCreate: device allocation failed");
	}
	*workspace = w;
	return ok();
}

CADR_API int cadrCullWorkspaceSetWaitLimit(cadrCullWorkspace* w, uint32_t microseconds)
{
	CADR_REQUIRE(w, "cadrCullWorkspaceSetWaitLimit: workspace is NULL");
	CADR_REQUIRE(microseconds > 0, "cadrCullWorkspaceSetWaitLimit: the limit must be positive");
	w->waitLimit = uint64_t(microseconds) * w->ticksPerUs;
	return ok();
}

CADR_API int cadrCullWorkspaceCheck(cadrCullWorkspace* w)
{
	CADR_REQUIRE(w, "cadrCullWorkspaceCheck: workspace is NULL");
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	CADR_REQUIRE(dev == w->device, "cadrCullWorkspaceCheck: workspace belongs to another device");
	CADR_CHECK_HIP(hipDeviceSynchronize(), "cadrCullWorkspaceCheck: hipDeviceSynchronize");
	uint32_t err = 0;
	CADR_CHECK_HIP(hipMemcpy(&err, w->ctl + kCtlError, sizeof(err), hipMemcpyDeviceToHost),
	               "cadrCullWorkspaceCheck: read");
	if (err == 0) return ok();
	const uint32_t zero = 0;
	CADR_CHECK_HIP(hipMemcpy(w->ctl + kCtlError, &zero, sizeof(zero), hipMemcpyHostToDevice),
	               "cadrCullWorkspaceCheck: clear");
	return setError(CADR_ERROR_TIMEOUT,
	                "cull: a look-back wait gave up (the tile list names a predecessor that no tile publishes); "
	                "the outputs of launch epoch " + std::to_string(err) + " are unspecified");
}

CADR_API int cadrCullWorkspaceDestroy(cadrCullWorkspace* w)
{
	if (!w) return ok();
	hipError_t e = hipDeviceSynchronize();
	(void)hipFree(w->status); (void)hipFree(w->ctl); (void)hipFree(w->firstWork); (void)hipFree(w->statusBase);
	delete w;
	CADR_CHECK_HIP(e, "cadrCullWorkspaceDestroy: hipDeviceSynchronize");
	return ok();
}

namespace {

// Leaf level of a fused launch against its job: node i is drawable i.
int checkLeaf(const cadrTransformLevel& L, const cadrCullJob& J, const std::string& fn)
{
	CADR_REQUIRE(L.count == J.numDrawables, fn + ": the leaf level must hold the job's drawables");
	CADR_REQUIRE(L.count == 0 || (L.localMatrices && L.worldMatrices), fn + ": null matrix array");
	CADR_REQUIRE((L.localMatrices % 16) == 0 && (L.worldMatrices % 16) == 0 && (L.parentWorld % 16) == 0 &&
	                 (L.spheres % 16) == 0,
	             fn + ": matrix and sphere arrays must be 16-byte aligned");
	CADR_REQUIRE(L.parentWorld == 0 || (L.parentIndex && L.parentIndex % 4 == 0),
	             fn + ": parentIndex is NULL or misaligned for a non-root level");
	return CADR_SUCCESS;
}

// One persistent launch of a batched kernel over numJobs jobs; leaves: the
// fused form's leaf levels (device memory), else null.
int launchBatch(KernelFn fn, const KernelEntry& ke, const cadrCullConfig* config, uint64_t deviceJobs,
                const cadrCullJob* hostJobs, uint32_t numJobs, cadrCullWorkspace* ws, cadrStream stream,
                const cadrTransformLevel* leaves, const std::string& name)
{
	CADR_REQUIRE(ws, name + ": workspace is NULL");
	if (numJobs == 0) CADR_RETURN_NO_LAUNCH(stream);
	CADR_REQUIRE(deviceJobs && hostJobs, name + ": null job array");
	CADR_REQUIRE(deviceJobs % 16 == 0, name + ": deviceJobs must be 16-byte aligned");
	CADR_REQUIRE(numJobs <= ws->maxJobs, name + ": more jobs than the workspace was created for");
	// the persistent grid; unless CADR_CULL_WHOLE_TILES, a job's last grid
	// tiles are handed out per view group
	const int cu = computeUnitCount();
	uint32_t perCU = config->workgroupsPerCU;
	if (perCU == 0) perCU = ke.block >= 512 ? 2 : 4;
	uint64_t grid = uint64_t(cu > 0 ? cu : 256) * perCU;
	const uint32_t splitTiles = (config->flags & CADR_CULL_WHOLE_TILES) ? 0u : uint32_t(std::min<uint64_t>(grid, 1u << 20));
	uint64_t totalWork = 0, statusNeed = 0;
	for (uint32_t j = 0; j < numJobs; j++) {
		const int rc = checkJob(hostJobs[j]);
		if (rc != CADR_SUCCESS) return rc;
		totalWork += jobWorkItems(hostJobs[j].numTiles, hostJobs[j].numViews, splitTiles);
		statusNeed += uint64_t(hostJobs[j].numViews) * hostJobs[j].numTiles;
	}
	CADR_REQUIRE(totalWork < 0xffffffffull, name + ": too many tiles in one launch");
	CADR_REQUIRE(statusNeed <= ws->statusWords, name + ": workspace has too few status words");
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	CADR_REQUIRE(dev == ws->device, name + ": workspace belongs to another device");

	hipStream_t s = hipStream_t(stream);
	const auto* jobs = reinterpret_cast<const cadrCullJob*>(deviceJobs);
	hipLaunchKernelGGL(cullSetupKernel, dim3(1), dim3(256), 0, s, jobs, numJobs, splitTiles, ws->firstWork,
	                   ws->statusBase, ws->ctl);
	CADR_CHECK_HIP(hipGetLastError(), "cullSetupKernel launch");
	if (totalWork == 0) CADR_RETURN_NO_LAUNCH(s);
	if (grid > totalWork) grid = totalWork;
	CullParams p{};
	p.jobs = jobs;
	p.firstWork = ws->firstWork;
	p.statusBase = ws->statusBase;
	p.status = ws->status;
	p.ctl = ws->ctl;
	p.numJobs = numJobs;
	p.totalWork = uint32_t(totalWork);
	p.splitTiles = splitTiles;
	p.flags = config->flags;
	p.waitLimit = ws->waitLimit;
	p.leaves = leaves;
	launchKernel(fn, dim3(uint32_t(grid)), dim3(ke.block), s, p);
	CADR_CHECK_HIP(hipGetLastError(), "cullKernel launch");
	return ok();
}

}  // namespace

CADR_API int cadrCullBatch(const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                           uint32_t numJobs, cadrCullWorkspace* ws, cadrStream stream)
{
	const KernelEntry* ke = findKernel(config);
	CADR_REQUIRE(ke && ke->batch, "cadrCullBatch: unsupported blockSize/itemsPerThread");
	return launchBatch(ke->batch, *ke, config, deviceJobs, hostJobs, numJobs, ws, stream, nullptr, "cadrCullBatch");
}

CADR_API int cadrPropagateCullBatch(uint64_t deviceLeaves, const cadrTransformLevel* hostLeaves,
                                    const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                                    uint32_t numJobs, cadrCullWorkspace* ws, cadrStream stream)
{
	const KernelEntry* ke = findKernel(config);
	CADR_REQUIRE(ke && ke->fusedBatch,
	             "cadrPropagateCullBatch: unsupported blockSize/itemsPerThread (256 x 4/8, 512 x 4)");
	if (numJobs == 0) CADR_RETURN_NO_LAUNCH(stream);
	CADR_REQUIRE(deviceLeaves && hostLeaves && hostJobs, "cadrPropagateCullBatch: null leaf or job array");
	CADR_REQUIRE(deviceLeaves % 8 == 0, "cadrPropagateCullBatch: deviceLeaves must be 8-byte aligned");
	for (uint32_t j = 0; j < numJobs; j++) {
		const int rc = checkLeaf(hostLeaves[j], hostJobs[j], "cadrPropagateCullBatch");
		if (rc != CADR_SUCCESS) return rc;
	}
	return launchBatch(ke->fusedBatch, *ke, config, deviceJobs, hostJobs, numJobs, ws, stream,
	                   reinterpret_cast<const cadrTransformLevel*>(deviceLeaves), "cadrPropagateCullBatch");
}

CADR_API int cadrCullView(const cadrCullConfig* config, uint64_t deviceJob, const cadrCullJob* hostJob,
                          cadrCullWorkspace* ws, cadrStream stream)
{
	CADR_REQUIRE(hostJob && hostJob->numViews == 1, "cadrCullView: the job must hold exactly one view");
	const KernelEntry* ke = findKernel(config);
	CADR_REQUIRE(ke, "cadrCullView: unsupported blockSize/itemsPerThread");
	if (!ke->view) return cadrCullBatch(config, deviceJob, hostJob, 1, ws, stream);
	CADR_REQUIRE(ws, "cadrCullView: workspace is NULL");
	CADR_REQUIRE(deviceJob && deviceJob % 16 == 0, "cadrCullView: deviceJob must be a 16-byte aligned address");
	const int rc = checkJob(*hostJob);
	if (rc != CADR_SUCCESS) return rc;
	CADR_REQUIRE(hostJob->numTiles <= ws->statusWords, "cadrCullView: workspace has too few status words");
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	CADR_REQUIRE(dev == ws->device, "cadrCullView: workspace belongs to another device");
	if (hostJob->numTiles == 0) CADR_RETURN_NO_LAUNCH(stream);
	CullParams p{};
	p.jobs = reinterpret_cast<const cadrCullJob*>(deviceJob);
	p.status = ws->status;
	p.ctl = ws->ctl;
	p.numJobs = 1;
	p.totalWork = hostJob->numTiles;
	p.flags = config->flags;
	p.waitLimit = ws->waitLimit;
	// a persistent kernel fills the GPU once; the other takes one tile per workgroup
	// more tiles than workgroups fit on the GPU at once: the persistent kernel
	// (grid = resident workgroups); else one tile per workgroup, which saves
	// the loop's extra ticket and barrier on a launch-bound view
	const int resident = ke->viewLoop ? residentWorkgroups(ke->viewLoop, ke->block) : 0;
	if (resident > 0 && hostJob->numTiles > uint32_t(resident))
		launchKernel(ke->viewLoop, dim3(uint32_t(resident)), dim3(ke->block), hipStream_t(stream), p);
	else
		launchKernel(ke->view, dim3(hostJob->numTiles), dim3(ke->block), hipStream_t(stream), p);
	CADR_CHECK_HIP(hipGetLastError(), "cullViewKernel launch");
	return ok();
}

CADR_API int cadrPropagateCullView(const cadrTransformLevel* leaf, const cadrCullConfig* config, uint64_t deviceJob,
                                   const cadrCullJob* hostJob, cadrCullWorkspace* ws, cadrStream stream)
{
	CADR_REQUIRE(leaf && hostJob, "cadrPropagateCullView: null argument");
	CADR_REQUIRE(hostJob->numViews == 1, "cadrPropagateCullView: the job must hold exactly one view");
	const KernelEntry* ke = findKernel(config);
	CADR_REQUIRE(ke && ke->fused,
	             "cadrPropagateCullView: unsupported blockSize/itemsPerThread (256 x 2/4/8/16, 512 x 2/4, 1024 x 1/2)");
	CADR_REQUIRE(ws, "cadrPropagateCullView: workspace is NULL");
	CADR_REQUIRE(deviceJob && deviceJob % 16 == 0, "cadrPropagateCullView: deviceJob must be a 16-byte aligned address");
	int rc = checkLeaf(*leaf, *hostJob, "cadrPropagateCullView");
	if (rc != CADR_SUCCESS) return rc;
	rc = checkJob(*hostJob);
	if (rc != CADR_SUCCESS) return rc;
	CADR_REQUIRE(hostJob->numTiles <= ws->statusWords, "cadrPropagateCullView: workspace has too few status words");
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	CADR_REQUIRE(dev == ws->device, "cadrPropagateCullView: workspace belongs to another device");
	if (hostJob->numTiles == 0) CADR_RETURN_NO_LAUNCH(stream);
	CullParams p{};
	p.jobs = reinterpret_cast<const cadrCullJob*>(deviceJob);
	p.status = ws->status;
	p.ctl = ws->ctl;
	p.numJobs = 1;
	p.totalWork = hostJob->numTiles;
	p.flags = config->flags;
	p.waitLimit = ws->waitLimit;
	p.leaf = LeafParams{leaf->localMatrices, leaf->parentIndex, leaf->parentWorld, leaf->worldMatrices, leaf->spheres,
	                    leaf->parentWorld ? leaf->parentCount : 0u};
	launchKernel(ke->fused, dim3(hostJob->numTiles), dim3(ke->block), hipStream_t(stream), p);
	CADR_CHECK_HIP(hipGetLastError(), "fused cullViewKernel launch");
	return ok();
}

// ------------------------------------------------ allocation-free variant
//
// cadrCullBatchNoAlloc: cadrCullBatch for callers that cannot allocate device
// memory.  Its workspace (look-back status words, control words, per-job
// offsets) is one file-scope device array, `scratch`, of fixed size; each
// device holds its own copy until the process ends.  Host code never passes
// the array's name to a kernel: the name is a host-side stand-in, and the
// device address is taken once per device with hipGetSymbolAddress.
namespace {

constexpr uint64_t kScratchStatusWords = 1ull << 21;  // 16 MiB: sum of numViews x numTiles per launch
constexpr uint32_t kScratchMaxJobs = 4096;
// layout in 64-bit words: status | ctl (kCtlWords uint32) | firstWork (jobs+1 uint32) | statusBase (jobs uint64)
constexpr uint64_t kScratchCtl = kScratchStatusWords;
constexpr uint64_t kScratchFirstWork = kScratchCtl + kCtlWords / 2;
constexpr uint64_t kScratchStatusBase = kScratchFirstWork + (uint64_t(kScratchMaxJobs) + 2) / 2;
constexpr uint64_t kScratchWords = kScratchStatusBase + kScratchMaxJobs;
static_assert(kCtlWords % 2 == 0, "ctl words fill whole 64-bit words");

}  // namespace

static __device__ __align__(256) uint64_t scratch[kScratchWords];

namespace {

struct ScratchDevice {
	uint64_t* ptr = nullptr;   // device address of `scratch` on this device
	uint64_t ticksPerUs = 100;
};

std::mutex g_scratchMutex;
std::vector<ScratchDevice> g_scratchDevices;

// Device address of this device's scratch, set up on the first call on the
// device: the address from hipGetSymbolAddress, status and control words
// cleared by a fill kernel (not hipMemsetAsync, so a caller may capture
// later calls) and the epoch set to 1, complete before the call returns.
int scratchFor(int dev, hipStream_t stream, ScratchDevice& out)
{
	std::lock_guard<std::mutex> lock(g_scratchMutex);
	if (int(g_scratchDevices.size()) <= dev) g_scratchDevices.resize(size_t(dev) + 1);
	ScratchDevice& d = g_scratchDevices[size_t(dev)];
	if (!d.ptr) {
		void* p = nullptr;
		CADR_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "cadrCullBatchNoAlloc: hipGetSymbolAddress(scratch)");
		auto* base = static_cast<uint64_t*>(p);
		const uint64_t a = reinterpret_cast<uint64_t>(base);
		int rc = cadrFillU32(a, 0u, kScratchStatusBase * sizeof(uint64_t), stream);
		if (rc == CADR_SUCCESS) rc = cadrFillU32(a + kScratchCtl * 8 + kCtlEpoch * 4, 1u, 4, stream);
		if (rc != CADR_SUCCESS) return rc;
		CADR_CHECK_HIP(hipStreamSynchronize(stream), "cadrCullBatchNoAlloc: clearing scratch");
		int khz = 0;
		CADR_CHECK_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev), "hipDeviceGetAttribute");
		d.ticksPerUs = khz >= 1000 ? uint64_t(khz) / 1000u : 1u;
		d.ptr = base;
	}
	out = d;
	return ok();
}

}  // namespace

CADR_API int cadrCullBatchNoAlloc(const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                                  uint32_t numJobs, cadrStream stream)
{
	const KernelEntry* ke = findKernel(config);
	CADR_REQUIRE(ke && ke->batch, "cadrCullBatchNoAlloc: unsupported blockSize/itemsPerThread");
	if (numJobs == 0) CADR_RETURN_NO_LAUNCH(stream);
	CADR_REQUIRE(hostJobs, "cadrCullBatchNoAlloc: hostJobs is NULL");
	if (numJobs > kScratchMaxJobs)
		return setError(CADR_ERROR_INVALID_VALUE, "cadrCullBatchNoAlloc: " + std::to_string(numJobs) +
		                                              " jobs, more than the " + std::to_string(kScratchMaxJobs) +
		                                              " its fixed scratch holds");
	uint64_t statusNeed = 0;
	for (uint32_t j = 0; j < numJobs; j++) statusNeed += uint64_t(hostJobs[j].numViews) * hostJobs[j].numTiles;
	if (statusNeed > kScratchStatusWords)
		return setError(CADR_ERROR_INVALID_VALUE,
		                "cadrCullBatchNoAlloc: the jobs need " + std::to_string(statusNeed) +
		                    " status words (numViews x numTiles summed over jobs), more than the " +
		                    std::to_string(kScratchStatusWords) + " of its fixed scratch");
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	ScratchDevice sd;
	const int rc = scratchFor(dev, hipStream_t(stream), sd);
	if (rc != CADR_SUCCESS) return rc;
	cadrCullWorkspace w;  // a view of the scratch array, nothing allocated
	w.status = sd.ptr;
	w.statusWords = kScratchStatusWords;
	w.ctl = reinterpret_cast<uint32_t*>(sd.ptr + kScratchCtl);
	w.firstWork = reinterpret_cast<uint32_t*>(sd.ptr + kScratchFirstWork);
	w.statusBase = sd.ptr + kScratchStatusBase;
	w.maxJobs = kScratchMaxJobs;
	w.device = dev;
	w.ticksPerUs = sd.ticksPerUs;
	w.waitLimit = kDefaultWaitLimitUs * sd.ticksPerUs;
	return launchBatch(ke->batch, *ke, config, deviceJobs, hostJobs, numJobs, &w, stream, nullptr,
	                   "cadrCullBatchNoAlloc");
}
