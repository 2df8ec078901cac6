// Frustum cull, LOD selection and per-StateSet stream compaction.
//
// The reference has no cull: processDrawables.comp emits every drawable and
// StateSet::recordToCommandBuffer draws StateSet k's full range
// [base_k, base_k + n_k) of the indirect buffer (src/CadR/StateSet.cpp:198-268).
// This pass keeps that range layout and writes, per view, the visible
// drawables of StateSet k compacted to the front of its range plus a draw
// count per StateSet (the operands of vkCmdDrawIndexedIndirectCount).
//
// One kernel serves the plain single-view pass, the many-view batch and the
// uneven many-scene batch:
//   - Work items are tiles of at most blockSize*itemsPerThread drawables that
//     never cross a StateSet boundary, over all jobs of the launch.  A
//     persistent grid takes tiles from an atomic ticket in increasing order,
//     so a tile only ever waits on tiles that are already running.
//   - A tile's spheres and mesh indices are read once from HBM (16 + 4 B per
//     drawable, coalesced: wave w owns a contiguous run of 64*itemsPerThread
//     drawables, item j of lane l is drawable w*64*itemsPerThread + j*64 + l),
//     staged in LDS, and then tested against every view of the job, 64 views
//     at a time.
//   - Visibility of (view, drawable) becomes one bit of a wave64 ballot kept
//     in LDS; the per-view tile count is a scalar popcount.
//   - Wave 0 then runs the decoupled look-back of all 64 views at once, lane
//     v walking the status chain of view v: status words carry {epoch, state
//     (aggregate / inclusive prefix), value}, the first tile of a StateSet
//     publishes its inclusive prefix at once, so chains never cross
//     StateSets.
//   - Writes are compacted per wave through an LDS index list: lane l emits
//     the l-th visible drawable of the wave, so a wave stores a contiguous run
//     of 20-byte commands; the LOD is evaluated for visible drawables only.
//   - A conservative bounding sphere of each tile, tested per view with a
//     margin far above FP32 rounding, lets whole (tile, view) pairs skip the
//     per-drawable tests (all outside or all inside).
#include "internal.h"

#include <algorithm>
#include <vector>

namespace cadr {

constexpr uint32_t kViewGroup = 64;   // views whose look-backs share one wave
constexpr uint32_t kCtlTicket = 0;    // ctl word: next tile ticket
constexpr uint32_t kCtlEpoch = 32;    // ctl word: launch epoch (own 128-B line)
constexpr uint32_t kCtlWords = 64;

constexpr uint32_t kStateAggregate = 1;
constexpr uint32_t kStatePrefix = 2;
constexpr uint32_t kValueMask = 0x3fffffffu;

struct CullParams {
	const cadrCullJob* jobs;
	const uint32_t* firstWork;   // [numJobs+1] prefix of numTiles
	const uint64_t* statusBase;  // [numJobs] prefix of numViews*numTiles
	uint64_t* status;
	uint32_t* ctl;
	uint32_t numJobs;
	uint32_t totalWork;
	uint32_t flags;
};

// Per-view constants, loaded with scalar loads (the view index is uniform).
struct ViewC {
	float4 r0, r1, r2;
	float nearZ, farZ, sx0, sx1, sy0, sy1, lodScale, th0, th1, th2;
};

__device__ __forceinline__ ViewC loadView(const cadrView* __restrict__ views, uint32_t v)
{
	const float4* p = reinterpret_cast<const float4*>(views + v);
	ViewC c;
	c.r0 = p[0];
	c.r1 = p[1];
	c.r2 = p[2];
	const float4 a = p[3];  // nearZ, farZ, sideX[0], sideX[1]
	const float4 b = p[4];  // sideY[0], sideY[1], lodScale, lodThreshold[0]
	const float4 d = p[5];  // lodThreshold[1], lodThreshold[2], reserved...
	c.nearZ = a.x; c.farZ = a.y; c.sx0 = a.z; c.sx1 = a.w;
	c.sy0 = b.x; c.sy1 = b.y; c.lodScale = b.z; c.th0 = b.w;
	c.th1 = d.x; c.th2 = d.y;
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

__device__ __forceinline__ uint32_t selectLod(float4 s, const ViewC& v, uint32_t lodCount)
{
	const float vz = viewCoord(v.r2, s);
	const float p = s.w * v.lodScale;
	uint32_t lod = uint32_t(p < v.th0 * vz) + uint32_t(p < v.th1 * vz) + uint32_t(p < v.th2 * vz);
	const uint32_t maxLod = (lodCount >= 1u ? (lodCount <= 4u ? lodCount : 4u) : 1u) - 1u;
	return lod < maxLod ? lod : maxLod;
}

// Tile-vs-view class: 0 = test each drawable, 1 = all outside, 2 = all inside.
//
// ts is a sphere containing every drawable sphere of the tile
// (|c_i - C| + r_i <= R).  Each quantity the drawable test compares,
// v.z and g(p) = |v.x|*sideX[0] - v.z*sideX[1] (the larger of two plane
// distances), is Lipschitz in the sphere centre with a constant bounded by
// L below, so a drawable's quantity differs from the tile centre's by at most
// L*|c_i - C| <= L*(R - r_i).  With L >= 1 the tile is wholly outside a plane
// when the centre is outside by more than L*R, and wholly inside when inside
// by more than L*R.  The margin m, about 2^-14 of the magnitudes involved, is
// some thousand times the FP32 rounding of the per-drawable expressions, so
// the float test of every drawable agrees with the class.
__device__ __forceinline__ uint32_t classifyTile(float4 ts, const ViewC& v)
{
	const float vx = viewCoord(v.r0, ts);
	const float vy = viewCoord(v.r1, ts);
	const float vz = viewCoord(v.r2, ts);
	const float l0 = sqrtf((v.r0.x * v.r0.x + v.r0.y * v.r0.y) + v.r0.z * v.r0.z);
	const float l1 = sqrtf((v.r1.x * v.r1.x + v.r1.y * v.r1.y) + v.r1.z * v.r1.z);
	const float l2 = sqrtf((v.r2.x * v.r2.x + v.r2.y * v.r2.y) + v.r2.z * v.r2.z);
	float L = fmaxf(1.0f, l2);
	L = fmaxf(L, l0 * fabsf(v.sx0) + l2 * fabsf(v.sx1));
	L = fmaxf(L, l1 * fabsf(v.sy0) + l2 * fabsf(v.sy1));
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

// Exclusive prefix of tile t in one chain: walk back over the predecessors,
// adding aggregates until an inclusive prefix (every StateSet's first tile
// publishes one at once) is found.  Each predecessor holds an earlier ticket
// and is running or done, so every wait ends.
__device__ uint32_t lookBack(const uint64_t* chain, uint32_t t, uint32_t epoch)
{
	uint32_t excl = 0;
	uint32_t p = t;
	while (p > 0) {
		--p;
		uint64_t s;
		for (;;) {
			s = __hip_atomic_load(chain + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			if (uint32_t(s >> 32) == epoch && (uint32_t(s) >> 30) != 0u) break;
			__builtin_amdgcn_s_sleep(1);
		}
		excl += uint32_t(s) & kValueMask;
		if ((uint32_t(s) >> 30) == kStatePrefix) break;
	}
	return excl;
}

template <int BLOCK, int ITEMS>
struct CullShared {
	static constexpr int WAVES = BLOCK / 64;
	static constexpr int TILE = BLOCK * ITEMS;
	float4 sph[TILE];
	uint32_t mesh[TILE];
	uint64_t mask[kViewGroup][WAVES][ITEMS];
	uint32_t cnt[kViewGroup][WAVES];
	uint32_t excl[kViewGroup];
	uint16_t list[WAVES][64 * ITEMS];
	float red[WAVES][6];
	uint32_t bad[WAVES];
	uint32_t ticket;
};

template <int BLOCK, int ITEMS>
__global__ __launch_bounds__(BLOCK) void cullKernel(CullParams P)
{
	using S = CullShared<BLOCK, ITEMS>;
	constexpr int WAVES = S::WAVES;
	constexpr int WAVE_ITEMS = 64 * ITEMS;
	__shared__ S sh;

	const uint32_t tid = threadIdx.x;
	const uint32_t lane = laneId();
	const uint32_t wave = tid / 64;
	const uint32_t epoch = P.ctl[kCtlEpoch];
	uint32_t job = 0;

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
				if (P.firstWork[mid] <= w) lo = mid; else hi = mid - 1;
			}
			job = uniform(lo);
		}
		const cadrCullJob J = P.jobs[job];
		const uint32_t t = w - P.firstWork[job];
		const uint4 tileWord = reinterpret_cast<const uint4*>(J.tiles)[t];
		uint32_t start = tileWord.x, count = tileWord.y;
		const uint32_t segBase = tileWord.z;
		const uint32_t k = tileWord.w & CADR_TILE_STATESET_MASK;
		const bool first = (tileWord.w & CADR_TILE_FIRST) != 0u;
		const bool last = (tileWord.w & CADR_TILE_LAST) != 0u;
		if (start >= J.numDrawables) count = 0;
		else if (count > J.numDrawables - start) count = J.numDrawables - start;
		if (count > uint32_t(S::TILE)) count = S::TILE;

		// load the tile: wave-contiguous runs, staged in LDS, bounds for the tile sphere
		const float4* __restrict__ spheres = reinterpret_cast<const float4*>(J.spheres);
		const uint32_t* __restrict__ meshIndex = reinterpret_cast<const uint32_t*>(J.meshIndex);
		float4 sp[ITEMS];
		bool nonFinite = false;
		float mn[3] = {__builtin_inff(), __builtin_inff(), __builtin_inff()};
		float mx[3] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
#pragma unroll
		for (int j = 0; j < ITEMS; j++) {
			const uint32_t e = wave * WAVE_ITEMS + j * 64 + lane;
			float4 s = make_float4(0.f, 0.f, 0.f, -__builtin_inff());
			uint32_t m = 0xffffffffu;
			if (e < count) {
				s = spheres[start + e];
				m = meshIndex[start + e];
				if (m >= J.numMeshes) s.w = -__builtin_inff();  // invalid mesh: never visible
			}
			sp[j] = s;
			sh.sph[e] = s;
			sh.mesh[e] = m;
			// a sphere other than a finite one with r >= 0 or the empty sphere
			// (r = -inf, never visible) disables the tile sphere for this tile
			if (e < count && s.w != -__builtin_inff() &&
			    !(__builtin_isfinite(s.x) && __builtin_isfinite(s.y) && __builtin_isfinite(s.z) &&
			      __builtin_isfinite(s.w) && s.w >= 0.f))
				nonFinite = true;
			if (s.w >= 0.f) {
				mn[0] = fminf(mn[0], s.x - s.w); mx[0] = fmaxf(mx[0], s.x + s.w);
				mn[1] = fminf(mn[1], s.y - s.w); mx[1] = fmaxf(mx[1], s.y + s.w);
				mn[2] = fminf(mn[2], s.z - s.w); mx[2] = fmaxf(mx[2], s.z + s.w);
			}
		}
		const bool useTileSphere = (P.flags & CADR_CULL_NO_TILE_SPHERES) == 0u;
		float4 tileSphere = make_float4(0.f, 0.f, 0.f, -1.f);
		if (useTileSphere) {
#pragma unroll
			for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
				for (int a = 0; a < 3; a++) {
					mn[a] = fminf(mn[a], __shfl_xor(mn[a], off));
					mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off));
				}
			}
			const bool waveNonFinite = __ballot(nonFinite) != 0ull;
			if (lane == 0) {
				sh.bad[wave] = waveNonFinite ? 1u : 0u;
				sh.red[wave][0] = mn[0]; sh.red[wave][1] = mn[1]; sh.red[wave][2] = mn[2];
				sh.red[wave][3] = mx[0]; sh.red[wave][4] = mx[1]; sh.red[wave][5] = mx[2];
			}
		}
		__syncthreads();
		bool tileUsable = false;
		if (useTileSphere) {
#pragma unroll
			for (int q = 0; q < WAVES; q++) {
				for (int a = 0; a < 3; a++) {
					mn[a] = fminf(mn[a], sh.red[q][a]);
					mx[a] = fmaxf(mx[a], sh.red[q][3 + a]);
				}
			}
			// finite bounds only (an empty or NaN tile keeps per-drawable tests)
			const float hx = (mx[0] - mn[0]) * 0.5f, hy = (mx[1] - mn[1]) * 0.5f, hz = (mx[2] - mn[2]) * 0.5f;
			const float cx = (mn[0] + mx[0]) * 0.5f, cy = (mn[1] + mx[1]) * 0.5f, cz = (mn[2] + mx[2]) * 0.5f;
			const float R = sqrtf(hx * hx + hy * hy + hz * hz);
			uint32_t anyBad = 0;
#pragma unroll
			for (int q = 0; q < WAVES; q++) anyBad |= sh.bad[q];
			tileUsable = anyBad == 0u && __builtin_isfinite(R) && __builtin_isfinite(cx) &&
			             __builtin_isfinite(cy) && __builtin_isfinite(cz);
			// inflate against rounding of the centre and radius above
			const float slack = (fabsf(cx) + fabsf(cy) + fabsf(cz) + R) * (1.0f / 65536.0f);
			tileSphere = make_float4(cx, cy, cz, R + slack);
		}
		const bool emptyTile = count == 0;

		const cadrView* __restrict__ views = reinterpret_cast<const cadrView*>(J.views);
		uint64_t* chain0 = P.status + P.statusBase[job];
		const cadrMesh* __restrict__ meshes = reinterpret_cast<const cadrMesh*>(J.meshes);

		for (uint32_t g0 = 0; g0 < J.numViews; g0 += kViewGroup) {
			const uint32_t vg = uniform(min(kViewGroup, J.numViews - g0));

			// 1. visibility bits and per-wave counts
			for (uint32_t v = 0; v < vg; v++) {
				const ViewC V = loadView(views, g0 + v);
				uint32_t cls = 0;
				if (emptyTile) cls = 1;
				else if (tileUsable) cls = uniform(classifyTile(tileSphere, V));
				uint32_t c = 0;
#pragma unroll
				for (int j = 0; j < ITEMS; j++) {
					const uint32_t e = wave * WAVE_ITEMS + j * 64 + lane;
					bool vis;
					if (cls == 1u) vis = false;
					else if (cls == 2u) vis = (e < count) && (sp[j].w >= 0.f);
					else vis = (e < count) && sphereVisible(sp[j], V);
					const uint64_t m = __ballot(vis);
					if (lane == 0) sh.mask[v][wave][j] = m;
					c += uint32_t(__builtin_popcountll(m));
				}
				if (lane == 0) sh.cnt[v][wave] = c;
			}
			__syncthreads();

			// 2. look-back, one view per lane of wave 0
			if (wave == 0 && lane < vg) {
				uint32_t total = 0;
#pragma unroll
				for (int q = 0; q < WAVES; q++) total += sh.cnt[lane][q];
				uint64_t* chain = chain0 + uint64_t(g0 + lane) * J.numTiles;
				uint32_t excl = 0;
				if (first) {
					publish(chain + t, packStatus(epoch, kStatePrefix, total));
				} else {
					publish(chain + t, packStatus(epoch, kStateAggregate, total));
					excl = lookBack(chain, t, epoch);
					publish(chain + t, packStatus(epoch, kStatePrefix, excl + total));
				}
				sh.excl[lane] = excl;
				if (last && k < J.numStateSets)
					reinterpret_cast<uint32_t*>(J.counts)[uint64_t(g0 + lane) * J.numStateSets + k] = excl + total;
			}
			__syncthreads();

			// 3. compacted writes
			for (uint32_t v = 0; v < vg; v++) {
				uint32_t wavePos = sh.excl[v];
				for (uint32_t q = 0; q < wave; q++) wavePos += sh.cnt[v][q];
				wavePos = uniform(wavePos);
				uint32_t run = 0;
#pragma unroll
				for (int j = 0; j < ITEMS; j++) {
					const uint64_t m = sh.mask[v][wave][j];
					if ((m >> lane) & 1ull) sh.list[wave][run + maskRank(m)] = uint16_t(j * 64 + lane);
					run += uint32_t(__builtin_popcountll(m));
				}
				run = uniform(run);
				if (run == 0) continue;
				waveSync();
				const ViewC V = loadView(views, g0 + v);
				uint32_t* out = reinterpret_cast<uint32_t*>(J.commands) + uint64_t(g0 + v) * J.commandStride * 5ull;
				for (uint32_t r = lane; r < run; r += 64) {
					const uint32_t e = wave * WAVE_ITEMS + sh.list[wave][r];
					const float4 s = sh.sph[e];
					const uint32_t mi = sh.mesh[e];
					const uint32_t pos = segBase + wavePos + r;
					if (pos >= J.commandStride) continue;
					const cadrMesh* mesh = meshes + mi;
					const uint32_t lod = selectLod(s, V, mesh->lodCount);
					uint32_t* o = out + uint64_t(pos) * 5ull;
					__builtin_nontemporal_store(mesh->indexCount[lod], o + 0);
					__builtin_nontemporal_store(1u, o + 1);
					__builtin_nontemporal_store(mesh->firstIndex[lod], o + 2);
					__builtin_nontemporal_store(uint32_t(mesh->vertexOffset), o + 3);
					__builtin_nontemporal_store(start + e, o + 4);
				}
				waveSync();
			}
			__syncthreads();
		}
	}
}

// Job offsets for the launch (prefix sums over jobs), ticket reset and a new
// epoch; one block, before the cull kernel on the same stream.
__global__ __launch_bounds__(256) void cullSetupKernel(const cadrCullJob* __restrict__ jobs, uint32_t numJobs,
                                                       uint32_t* __restrict__ firstWork,
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
			nt = jobs[j].numTiles;
			ns = uint64_t(jobs[j].numViews) * nt;
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
};

using namespace cadr;

namespace {

using KernelFn = void (*)(CullParams);

struct KernelEntry {
	uint32_t block, items;
	KernelFn fn;
};

const KernelEntry kKernels[] = {
    {256, 2, cullKernel<256, 2>},  {256, 4, cullKernel<256, 4>}, {256, 8, cullKernel<256, 8>},
    {512, 2, cullKernel<512, 2>},  {512, 4, cullKernel<512, 4>}, {128, 8, cullKernel<128, 8>},
    {128, 16, cullKernel<128, 16>},
};

const KernelEntry* findKernel(const cadrCullConfig* c)
{
	if (!c) return nullptr;
	for (const KernelEntry& k : kKernels)
		if (k.block == c->blockSize && k.items == c->itemsPerThread) return &k;
	return nullptr;
}

}  // namespace

CADR_API uint32_t cadrCullTileSize(const cadrCullConfig* config)
{
	const KernelEntry* k = findKernel(config);
	return k ? k->block * k->items : 0u;
}

CADR_API int cadrMakeView(const float eye[3], const float dir[3], const float up[3], float fovy, float aspect,
                          float zNear, float zFar, float viewportHeight, const float* lodThresholds, cadrView* out);

CADR_API int cadrBuildCullTiles(const uint32_t* stateSetCounts, uint32_t numStateSets, uint32_t tileSize,
                                cadrCullTile* tiles, uint32_t capacity, uint32_t* numTiles)
{
	CADR_REQUIRE(numTiles, "cadrBuildCullTiles: numTiles is NULL");
	CADR_REQUIRE(tileSize > 0, "cadrBuildCullTiles: tileSize is 0");
	CADR_REQUIRE(numStateSets == 0 || stateSetCounts, "cadrBuildCullTiles: stateSetCounts is NULL");
	CADR_REQUIRE(numStateSets <= CADR_TILE_STATESET_MASK, "cadrBuildCullTiles: too many StateSets");
	uint64_t base = 0, n = 0;
	for (uint32_t k = 0; k < numStateSets; k++) {
		const uint64_t c = stateSetCounts[k];
		const uint64_t nt = c == 0 ? 1 : (c + tileSize - 1) / tileSize;
		for (uint64_t q = 0; q < nt; q++) {
			if (tiles) {
				if (n >= capacity) return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: capacity too small");
				const uint64_t s = base + q * tileSize;
				const uint64_t cnt = c == 0 ? 0 : std::min<uint64_t>(tileSize, c - q * tileSize);
				uint32_t flags = k;
				if (q == 0) flags |= CADR_TILE_FIRST;
				if (q + 1 == nt) flags |= CADR_TILE_LAST;
				tiles[n] = cadrCullTile{uint32_t(s), uint32_t(cnt), uint32_t(base), flags};
			}
			n++;
		}
		base += c;
		if (base > 0xffffffffull) return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: more than 2^32 drawables");
	}
	if (n > 0xffffffffull) return setError(CADR_ERROR_INVALID_VALUE, "cadrBuildCullTiles: too many tiles");
	*numTiles = uint32_t(n);
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

CADR_API int cadrCullWorkspaceDestroy(cadrCullWorkspace* w)
{
	if (!w) return ok();
	hipError_t e = hipDeviceSynchronize();
	(void)hipFree(w->status); (void)hipFree(w->ctl); (void)hipFree(w->firstWork); (void)hipFree(w->statusBase);
	delete w;
	CADR_CHECK_HIP(e, "cadrCullWorkspaceDestroy: hipDeviceSynchronize");
	return ok();
}

CADR_API int cadrCullBatch(const cadrCullConfig* config, uint64_t deviceJobs, const cadrCullJob* hostJobs,
                           uint32_t numJobs, cadrCullWorkspace* ws, cadrStream stream)
{
	const KernelEntry* ke = findKernel(config);
	CADR_REQUIRE(ke, "cadrCullBatch: unsupported blockSize/itemsPerThread");
	CADR_REQUIRE(ws, "cadrCullBatch: workspace is NULL");
	if (numJobs == 0) return ok();
	CADR_REQUIRE(deviceJobs && hostJobs, "cadrCullBatch: null job array");
	CADR_REQUIRE(deviceJobs % 16 == 0, "cadrCullBatch: deviceJobs must be 16-byte aligned");
	CADR_REQUIRE(numJobs <= ws->maxJobs, "cadrCullBatch: more jobs than the workspace was created for");
	const uint32_t tileSize = ke->block * ke->items;
	uint64_t totalWork = 0, statusNeed = 0;
	for (uint32_t j = 0; j < numJobs; j++) {
		const cadrCullJob& J = hostJobs[j];
		CADR_REQUIRE(J.numTiles == 0 || J.tiles, "cadrCullBatch: job without tiles");
		CADR_REQUIRE(J.numTiles == 0 || (J.spheres && J.meshIndex && J.meshes && J.views && J.commands && J.counts),
		             "cadrCullBatch: job with a null array");
		CADR_REQUIRE(J.spheres % 16 == 0 && J.meshes % 16 == 0 && J.views % 16 == 0 && J.tiles % 16 == 0 &&
		                 J.meshIndex % 4 == 0 && J.commands % 4 == 0 && J.counts % 4 == 0,
		             "cadrCullBatch: misaligned job array (spheres, meshes, views, tiles need 16 bytes)");
		CADR_REQUIRE(J.commandStride >= J.numDrawables, "cadrCullBatch: commandStride < numDrawables");
		CADR_REQUIRE(J.numTiles == 0 || J.numViews > 0 || true, "");
		totalWork += J.numTiles;
		statusNeed += uint64_t(J.numViews) * J.numTiles;
	}
	CADR_REQUIRE(totalWork < 0xffffffffull, "cadrCullBatch: too many tiles in one launch");
	CADR_REQUIRE(statusNeed <= ws->statusWords, "cadrCullBatch: workspace has too few status words");
	(void)tileSize;
	int dev = 0;
	CADR_CHECK_HIP(hipGetDevice(&dev), "hipGetDevice");
	CADR_REQUIRE(dev == ws->device, "cadrCullBatch: workspace belongs to another device");

	hipStream_t s = hipStream_t(stream);
	const auto* jobs = reinterpret_cast<const cadrCullJob*>(deviceJobs);
	hipLaunchKernelGGL(cullSetupKernel, dim3(1), dim3(256), 0, s, jobs, numJobs, ws->firstWork, ws->statusBase, ws->ctl);
	CADR_CHECK_HIP(hipGetLastError(), "cullSetupKernel launch");
	if (totalWork == 0) return ok();

	const int cu = computeUnitCount();
	uint32_t perCU = config->workgroupsPerCU;
	if (perCU == 0) perCU = ke->block >= 512 ? 2 : 4;
	uint64_t grid = uint64_t(cu > 0 ? cu : 256) * perCU;
	if (grid > totalWork) grid = totalWork;
	CullParams p;
	p.jobs = jobs;
	p.firstWork = ws->firstWork;
	p.statusBase = ws->statusBase;
	p.status = ws->status;
	p.ctl = ws->ctl;
	p.numJobs = numJobs;
	p.totalWork = uint32_t(totalWork);
	p.flags = config->flags;
	hipLaunchKernelGGL(ke->fn, dim3(uint32_t(grid)), dim3(ke->block), 0, s, p);
	CADR_CHECK_HIP(hipGetLastError(), "cullKernel launch");
	return ok();
}

CADR_API int cadrCullView(const cadrCullConfig* config, uint64_t deviceJob, const cadrCullJob* hostJob,
                          cadrCullWorkspace* workspace, cadrStream stream)
{
	CADR_REQUIRE(hostJob && hostJob->numViews == 1, "cadrCullView: the job must hold exactly one view");
	return cadrCullBatch(config, deviceJob, hostJob, 1, workspace, stream);
}
