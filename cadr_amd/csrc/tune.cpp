// Autotuner of the cull launch shape (cadrCullAutotune).
//
// Built on the public C ABI only: tiles per candidate tile size
// (cadrBuildCullTiles), one workspace sized for the largest tiling, real
// launches timed with HIP events on the caller's stream.  Each candidate's
// per-view counts are compared with the first candidate's, so a tuning run
// doubles as a consistency check of the kernel shapes.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <vector>

#include "internal.h"

using namespace cadr;

namespace {

struct Tiling {
	std::vector<void*> tileBuffers;
	void* deviceJobs = nullptr;
	std::vector<cadrCullJob> hostJobs;
	uint64_t statusWords = 0;
};

void freeTiling(Tiling& t)
{
	for (void* p : t.tileBuffers) (void)hipFree(p);
	if (t.deviceJobs) (void)hipFree(t.deviceJobs);
	t.tileBuffers.clear();
	t.deviceJobs = nullptr;
}

int buildTiling(const cadrCullTuneJob* jobs, uint32_t numJobs, uint32_t tileSize, Tiling& t)
{
	t.hostJobs.resize(numJobs);
	for (uint32_t j = 0; j < numJobs; j++) {
		const cadrCullJob& J = jobs[j].job;
		uint32_t n = 0;
		int rc = cadrBuildCullTiles(jobs[j].stateSetCounts, J.numStateSets, tileSize, nullptr, 0, &n);
		if (rc != CADR_SUCCESS) return rc;
		std::vector<cadrCullTile> tiles(n);
		rc = cadrBuildCullTiles(jobs[j].stateSetCounts, J.numStateSets, tileSize, tiles.data(), n, &n);
		if (rc != CADR_SUCCESS) return rc;
		void* d = nullptr;
		CADR_CHECK_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(cadrCullTile)), "hipMalloc(tiles)");
		t.tileBuffers.push_back(d);
		CADR_CHECK_HIP(hipMemcpy(d, tiles.data(), n * sizeof(cadrCullTile), hipMemcpyHostToDevice), "hipMemcpy(tiles)");
		t.hostJobs[j] = J;
		t.hostJobs[j].tiles = reinterpret_cast<uint64_t>(d);
		t.hostJobs[j].numTiles = n;
		t.statusWords += uint64_t(n) * J.numViews;
	}
	CADR_CHECK_HIP(hipMalloc(&t.deviceJobs, numJobs * sizeof(cadrCullJob)), "hipMalloc(jobs)");
	CADR_CHECK_HIP(hipMemcpy(t.deviceJobs, t.hostJobs.data(), numJobs * sizeof(cadrCullJob), hipMemcpyHostToDevice),
	               "hipMemcpy(jobs)");
	return CADR_SUCCESS;
}

int launch(const cadrCullConfig& c, bool single, Tiling& t, cadrCullWorkspace* ws, cadrStream stream)
{
	const uint64_t dj = reinterpret_cast<uint64_t>(t.deviceJobs);
	if (single) return cadrCullView(&c, dj, t.hostJobs.data(), ws, stream);
	return cadrCullBatch(&c, dj, t.hostJobs.data(), uint32_t(t.hostJobs.size()), ws, stream);
}

int readCounts(const std::vector<cadrCullJob>& jobs, std::vector<uint32_t>& out, cadrStream stream)
{
	out.clear();
	for (const cadrCullJob& J : jobs) {
		const size_t at = out.size(), n = size_t(J.numViews) * J.numStateSets;
		out.resize(at + n);
		CADR_CHECK_HIP(hipMemcpyAsync(out.data() + at, reinterpret_cast<const void*>(J.counts), n * 4,
		                              hipMemcpyDeviceToHost, reinterpret_cast<hipStream_t>(stream)),
		               "hipMemcpyAsync(counts)");
	}
	CADR_CHECK_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)), "hipStreamSynchronize");
	return CADR_SUCCESS;
}

}  // namespace

CADR_API int cadrCullAutotune(const cadrCullTuneJob* jobs, uint32_t numJobs, uint32_t flags, uint32_t reps,
                              const cadrCullConfig* candidates, uint32_t numCandidates,
                              cadrCullTuneResult* results, uint32_t resultCapacity, uint32_t* numResults,
                              cadrCullConfig* best, cadrStream stream)
{
	CADR_REQUIRE(jobs && numJobs > 0 && best, "cadrCullAutotune: no jobs or best is NULL");
	CADR_REQUIRE(!results || numResults, "cadrCullAutotune: numResults is NULL");
	const bool single = (flags & CADR_TUNE_SINGLE_VIEW) != 0;
	CADR_REQUIRE(!single || (numJobs == 1 && jobs[0].job.numViews == 1),
	             "cadrCullAutotune: single-view tuning needs one job with one view");
	for (uint32_t j = 0; j < numJobs; j++)
		CADR_REQUIRE(jobs[j].job.numStateSets == 0 || jobs[j].stateSetCounts, "cadrCullAutotune: stateSetCounts is NULL");
	if (reps == 0) reps = 5;

	std::vector<cadrCullConfig> cands;
	if (candidates && numCandidates) {
		cands.assign(candidates, candidates + numCandidates);
	} else {
		for (uint32_t block : {128u, 256u, 512u, 1024u})
			for (uint32_t items : {1u, 2u, 4u, 8u, 16u}) {
				const cadrCullConfig c{block, items, 0, 0};
				if (!cullHasKernel(&c, single ? CullPath::View : CullPath::Batch)) continue;
				cands.push_back(c);
				if (!single) cands.push_back(cadrCullConfig{block, items, 0, CADR_CULL_NO_TILE_SPHERES});
			}
	}
	for (const cadrCullConfig& c : cands) CADR_REQUIRE(cadrCullTileSize(&c) > 0, "cadrCullAutotune: unsupported candidate");

	// one tiling per distinct tile size
	std::map<uint32_t, Tiling> tilings;
	int rc = CADR_SUCCESS;
	uint64_t maxStatus = 1;
	for (const cadrCullConfig& c : cands) {
		const uint32_t ts = cadrCullTileSize(&c);
		if (tilings.count(ts)) continue;
		rc = buildTiling(jobs, numJobs, ts, tilings[ts]);
		if (rc != CADR_SUCCESS) break;
		maxStatus = std::max(maxStatus, tilings[ts].statusWords);
	}
	cadrCullWorkspace* ws = nullptr;
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (rc == CADR_SUCCESS) rc = cadrCullWorkspaceCreate(maxStatus, numJobs, &ws);
	if (rc == CADR_SUCCESS && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
		rc = setError(CADR_ERROR_DEVICE, "cadrCullAutotune: hipEventCreate failed");

	const hipStream_t hs = reinterpret_cast<hipStream_t>(stream);
	std::vector<uint32_t> firstCounts, counts;
	bool haveFirst = false;
	float bestMs = std::numeric_limits<float>::infinity();
	cadrCullConfig bestCfg = cands.empty() ? cadrCullConfig{256, 4, 0, 0} : cands[0];
	uint32_t nres = 0;
	for (size_t i = 0; rc == CADR_SUCCESS && i < cands.size(); i++) {
		const cadrCullConfig& c = cands[i];
		Tiling& t = tilings[cadrCullTileSize(&c)];
		float ms = std::numeric_limits<float>::infinity();
		uint32_t agrees = 0;
		bool okc = launch(c, single, t, ws, stream) == CADR_SUCCESS && hipEventRecord(e0, hs) == hipSuccess;
		for (uint32_t r = 0; okc && r < reps; r++) okc = launch(c, single, t, ws, stream) == CADR_SUCCESS;
		okc = okc && hipEventRecord(e1, hs) == hipSuccess && hipEventSynchronize(e1) == hipSuccess;
		float total = 0.f;
		if (okc && hipEventElapsedTime(&total, e0, e1) == hipSuccess) {
			ms = total / float(reps);
			rc = readCounts(t.hostJobs, counts, stream);
			if (rc != CADR_SUCCESS) break;
			if (!haveFirst) {
				firstCounts = counts;
				haveFirst = true;
			}
			agrees = counts == firstCounts ? 1u : 0u;
			if (agrees && ms < bestMs) {
				bestMs = ms;
				bestCfg = c;
			}
		} else {
			(void)hipGetLastError();
		}
		if (results && nres < resultCapacity) results[nres++] = cadrCullTuneResult{c, ms, agrees};
	}
	if (numResults) *numResults = nres;
	if (e0) (void)hipEventDestroy(e0);
	if (e1) (void)hipEventDestroy(e1);
	if (ws) cadrCullWorkspaceDestroy(ws);
	(void)hipStreamSynchronize(hs);
	for (auto& kv : tilings) freeTiling(kv.second);
	if (rc != CADR_SUCCESS) return rc;
	if (!std::isfinite(bestMs)) return setError(CADR_ERROR_DEVICE, "cadrCullAutotune: no candidate ran");
	*best = bestCfg;
	return CADR_SUCCESS;
}
