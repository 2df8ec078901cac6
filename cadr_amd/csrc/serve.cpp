// Serving loop: resident scenes, cull requests admitted by output memory and
// batch slots, continuous batching across scenes, latency accounting.
//
// The reference has no serving mode (it renders one scene per frame); this is
// the many-view form of the same pass.  A scheduler thread owns the device
// side:
//   - retire: a slot whose launch has finished (its event completed) hands
//     its requests' per-StateSet counts, already copied to pinned host memory
//     by the same stream, to the completion queue;
//   - launch: when a slot is free and requests wait, it takes them in arrival
//     order up to maxBatchViews, groups them by scene (one cull job per scene,
//     the scene read once for all its views), allocates each job's output
//     block from the output pool (admission by memory: the head of the queue
//     waits until its block fits), uploads views and job descriptors from the
//     slot's pinned buffers and issues cadrCullBatch plus the counts copy on
//     the slot's stream.
// Launches on different slots run concurrently; requests that arrive while
// launches run are taken by the next free slot.  Each request's commands and
// counts are what cadrCullView computes for it alone: the kernel's result for
// one (view, drawable) does not depend on the other views or jobs of a launch.
#include "internal.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

uint64_t nowNs()
{
	return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
	                    std::chrono::steady_clock::now().time_since_epoch())
	                    .count());
}

// First-fit allocator over one device allocation, 256-byte granules.
class OutputPool {
public:
	~OutputPool()
	{
		if (_base) (void)hipFree(reinterpret_cast<void*>(_base));
	}
	int init(uint64_t bytes)
	{
		void* p = nullptr;
		if (hipMalloc(&p, bytes) != hipSuccess) {
			(void)hipGetLastError();
			return cadr::setError(CADR_ERROR_OUT_OF_RESOURCES, "serve: cannot allocate the output budget");
		}
		_base = reinterpret_cast<uint64_t>(p);
		_size = bytes;
		_free[0] = bytes;
		return CADR_SUCCESS;
	}
	static uint64_t round(uint64_t b) { return (b + 255) & ~uint64_t(255); }
	bool alloc(uint64_t bytes, uint64_t& addr)
	{
		bytes = round(bytes);
		for (auto it = _free.begin(); it != _free.end(); ++it) {
			if (it->second >= bytes) {
				addr = _base + it->first;
				const uint64_t off = it->first + bytes, rest = it->second - bytes;
				_free.erase(it);
				if (rest) _free[off] = rest;
				_used += bytes;
				return true;
			}
		}
		return false;
	}
	void release(uint64_t addr, uint64_t bytes)
	{
		bytes = round(bytes);
		uint64_t off = addr - _base;
		_used -= bytes;
		auto next = _free.lower_bound(off);
		if (next != _free.end() && off + bytes == next->first) {
			bytes += next->second;
			next = _free.erase(next);
		}
		if (next != _free.begin()) {
			auto prev = std::prev(next);
			if (prev->first + prev->second == off) {
				prev->second += bytes;
				return;
			}
		}
		_free[off] = bytes;
	}
	uint64_t size() const { return _size; }
	uint64_t freeBytes() const { return _size - _used; }

private:
	uint64_t _base = 0, _size = 0, _used = 0;
	std::map<uint64_t, uint64_t> _free;  // offset -> bytes
};

struct Request {
	uint64_t id = 0;
	uint32_t scene = 0;
	cadrView view{};
	uint64_t submitNs = 0, completeNs = 0;
	uint64_t commands = 0, counts = 0;
	uint32_t block = 0;
	std::vector<uint32_t> hostCounts;
	uint64_t visible = 0;
	uint32_t tag = 0;
	int status = CADR_SUCCESS;
	bool done = false;
};

struct OutputBlock {
	uint64_t addr = 0, bytes = 0;
	uint32_t refs = 0;
};

struct Slot {
	hipStream_t stream = nullptr;
	hipEvent_t done = nullptr;
	cadrCullWorkspace* ws = nullptr;
	uint64_t dViews = 0, dJobs = 0;
	cadrView* hViews = nullptr;
	cadrCullJob* hJobs = nullptr;
	uint32_t* hCounts = nullptr;
	std::vector<Request*> reqs;
	std::vector<uint64_t> countOffset;  // per request, in uint32 units of hCounts
	bool busy = false;
};

struct SceneRec {
	cadrServerScene desc{};
	std::vector<uint32_t> counts;
	uint64_t tiles = 0;
	uint32_t numTiles = 0;
	uint64_t bytesPerView() const
	{
		return OutputPool::round(uint64_t(desc.numDrawables) * 20u) + OutputPool::round(uint64_t(desc.numStateSets) * 4u);
	}
};

}  // namespace

struct cadrServer {
	cadrServerConfig cfg{};
	uint32_t tileSize = 0;
	int device = 0;
	uint32_t maxStateSets = 1, maxTiles = 1;
	std::vector<SceneRec> scenes;
	std::vector<Slot> slots;
	OutputPool pool;
	std::vector<OutputBlock> blocks;
	std::vector<uint32_t> freeBlocks;
	std::mutex m;
	std::condition_variable cvWork, cvDone;
	std::deque<Request*> queue;
	std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
	std::deque<uint64_t> completed;
	std::atomic<bool> stop{false};
	std::thread worker;
	bool started = false;
	uint64_t nextId = 1;
	uint64_t inFlight = 0;      // requests queued or running
	uint64_t wake = 0;          // bumped by submissions and releases
	uint64_t wakeSeen = 0;      // value of wake at the last launch attempt
	std::atomic<uint64_t> launches{0}, launchedViews{0};
	std::atomic<uint64_t> sceneBytesLaunched{0};  // scene arrays each launch reads once per job
	std::string workerMessage;  // first device failure of a launch (guarded by m)

	~cadrServer();
	void run();
	bool retire();
	bool launch(Slot& slot);
	// completes the requests of a slot (status != 0: failed, counts zero); called with m held
	void complete(Slot& slot, int status, uint64_t t);
	void recordFailure(const std::string& what);
	int submitAt(uint32_t scene, const cadrView* view, uint64_t arrivalNs, uint32_t tag, uint64_t* id);
};

using namespace cadr;

cadrServer::~cadrServer()
{
	stop = true;
	cvWork.notify_all();
	if (worker.joinable()) worker.join();
	(void)hipSetDevice(device);
	(void)hipDeviceSynchronize();
	for (Slot& s : slots) {
		if (s.stream) (void)hipStreamDestroy(s.stream);
		if (s.done) (void)hipEventDestroy(s.done);
		if (s.ws) cadrCullWorkspaceDestroy(s.ws);
		(void)hipFree(reinterpret_cast<void*>(s.dViews));
		(void)hipFree(reinterpret_cast<void*>(s.dJobs));
		if (s.hViews) (void)hipHostFree(s.hViews);
		if (s.hJobs) (void)hipHostFree(s.hJobs);
		if (s.hCounts) (void)hipHostFree(s.hCounts);
	}
	for (SceneRec& sc : scenes) (void)hipFree(reinterpret_cast<void*>(sc.tiles));
}

void cadrServer::complete(Slot& s, int status, uint64_t t)
{
	for (size_t r = 0; r < s.reqs.size(); r++) {
		Request* q = s.reqs[r];
		const uint32_t K = scenes[q->scene].desc.numStateSets;
		if (status == CADR_SUCCESS)
			q->hostCounts.assign(s.hCounts + s.countOffset[r], s.hCounts + s.countOffset[r] + K);
		else
			q->hostCounts.assign(K, 0u);
		q->visible = 0;
		for (uint32_t c : q->hostCounts) q->visible += c;
		q->status = status;
		q->completeNs = t;
		q->done = true;
		completed.push_back(q->id);
		inFlight--;
	}
	s.reqs.clear();
	s.busy = false;
}

void cadrServer::recordFailure(const std::string& what)
{
	std::lock_guard<std::mutex> lock(m);
	if (workerMessage.empty()) workerMessage = what;
}

// Moves finished launches' requests to the completion queue.
bool cadrServer::retire()
{
	bool any = false;
	for (Slot& s : slots) {
		if (!s.busy) continue;
		const hipError_t e = hipEventQuery(s.done);
		if (e == hipErrorNotReady) continue;
		if (e != hipSuccess) recordFailure(std::string("serve: launch failed on the device: ") + hipGetErrorString(e));
		{
			std::lock_guard<std::mutex> lock(m);
			complete(s, e == hipSuccess ? CADR_SUCCESS : CADR_ERROR_DEVICE, nowNs());
		}
		any = true;
		cvDone.notify_all();
	}
	return any;
}

// Takes waiting requests for one launch on a free slot.
bool cadrServer::launch(Slot& s)
{
	std::vector<Request*> take;
	std::vector<uint32_t> jobScene;
	std::vector<std::vector<Request*>> jobReqs;
	{
		std::lock_guard<std::mutex> lock(m);
		wakeSeen = wake;
		if (queue.empty()) return false;
		std::unordered_map<uint32_t, uint32_t> jobOf;
		uint64_t budget = pool.freeBytes();
		while (!queue.empty() && take.size() < cfg.maxBatchViews) {
			Request* q = queue.front();
			const uint64_t need = scenes[q->scene].bytesPerView();
			if (need > budget) break;  // admission by memory: the head waits for outputs to be released
			budget -= need;
			queue.pop_front();
			take.push_back(q);
			auto it = jobOf.find(q->scene);
			if (it == jobOf.end()) {
				jobOf[q->scene] = uint32_t(jobScene.size());
				jobScene.push_back(q->scene);
				jobReqs.emplace_back();
				jobReqs.back().push_back(q);
			} else {
				jobReqs[it->second].push_back(q);
			}
		}
		if (take.empty()) return false;

		// one output block per job: commands of every view, then counts of every view
		for (size_t j = 0; j < jobScene.size(); j++) {
			const SceneRec& sc = scenes[jobScene[j]];
			const uint64_t k = jobReqs[j].size();
			const uint64_t cmdBytes = OutputPool::round(uint64_t(sc.desc.numDrawables) * 20u * k);
			const uint64_t bytes = cmdBytes + OutputPool::round(uint64_t(sc.desc.numStateSets) * 4u * k);
			uint64_t addr = 0;
			if (!pool.alloc(bytes, addr)) {
				// fragmentation: this job and the rest go back to the front, in order
				for (size_t r = j; r < jobScene.size(); r++) {
					for (Request* q : jobReqs[r]) take.erase(std::find(take.begin(), take.end(), q));
				}
				std::vector<Request*> back;
				for (size_t r = j; r < jobScene.size(); r++)
					for (Request* q : jobReqs[r]) back.push_back(q);
				std::sort(back.begin(), back.end(), [](Request* a, Request* b) { return a->id < b->id; });
				for (auto it = back.rbegin(); it != back.rend(); ++it) queue.push_front(*it);
				jobScene.resize(j);
				jobReqs.resize(j);
				break;
			}
			uint32_t bi;
			if (!freeBlocks.empty()) {
				bi = freeBlocks.back();
				freeBlocks.pop_back();
			} else {
				bi = uint32_t(blocks.size());
				blocks.emplace_back();
			}
			blocks[bi] = OutputBlock{addr, bytes, uint32_t(k)};
			for (uint64_t v = 0; v < k; v++) {
				Request* q = jobReqs[j][v];
				q->block = bi;
				q->commands = addr + v * uint64_t(sc.desc.numDrawables) * 20u;
				q->counts = addr + cmdBytes + v * uint64_t(sc.desc.numStateSets) * 4u;
			}
		}
		if (jobScene.empty()) return false;
	}

	// descriptors in the slot's pinned buffers (the slot is idle, so they are free)
	uint32_t nv = 0;
	s.reqs.clear();
	s.countOffset.clear();
	uint64_t countsAt = 0;
	for (size_t j = 0; j < jobScene.size(); j++) {
		const SceneRec& sc = scenes[jobScene[j]];
		cadrCullJob& J = s.hJobs[j];
		J.spheres = sc.desc.spheres;
		J.meshIndex = sc.desc.meshIndex;
		J.meshes = sc.desc.meshes;
		J.tiles = sc.tiles;
		J.views = s.dViews + uint64_t(nv) * sizeof(cadrView);
		J.commands = jobReqs[j][0]->commands;
		J.counts = jobReqs[j][0]->counts;
		J.numDrawables = sc.desc.numDrawables;
		J.numTiles = sc.numTiles;
		J.numStateSets = sc.desc.numStateSets;
		J.numViews = uint32_t(jobReqs[j].size());
		J.commandStride = sc.desc.numDrawables;
		J.numMeshes = sc.desc.numMeshes;
		for (Request* q : jobReqs[j]) {
			s.hViews[nv++] = q->view;
			s.reqs.push_back(q);
			s.countOffset.push_back(countsAt);
			countsAt += sc.desc.numStateSets;
		}
	}
	const uint32_t numJobs = uint32_t(jobScene.size());
	bool okLaunch = hipMemcpyAsync(reinterpret_cast<void*>(s.dViews), s.hViews, uint64_t(nv) * sizeof(cadrView),
	                               hipMemcpyHostToDevice, s.stream) == hipSuccess &&
	                hipMemcpyAsync(reinterpret_cast<void*>(s.dJobs), s.hJobs, uint64_t(numJobs) * sizeof(cadrCullJob),
	                               hipMemcpyHostToDevice, s.stream) == hipSuccess;
	okLaunch = okLaunch && cadrCullBatch(&cfg.cull, s.dJobs, s.hJobs, numJobs, s.ws, s.stream) == CADR_SUCCESS;
	uint64_t at = 0;
	for (size_t j = 0; j < jobScene.size() && okLaunch; j++) {
		const uint64_t bytes = uint64_t(s.hJobs[j].numViews) * s.hJobs[j].numStateSets * 4u;
		okLaunch = hipMemcpyAsync(s.hCounts + at, reinterpret_cast<const void*>(s.hJobs[j].counts), bytes,
		                          hipMemcpyDeviceToHost, s.stream) == hipSuccess;
		at += bytes / 4u;
	}
	okLaunch = okLaunch && hipEventRecord(s.done, s.stream) == hipSuccess;
	launches++;
	launchedViews += nv;
	{
		uint64_t sceneBytes = 0;  // spheres (16 B) and mesh indices (4 B) of every job's drawables
		for (uint32_t j = 0; j < numJobs; j++) sceneBytes += uint64_t(s.hJobs[j].numDrawables) * 20u;
		sceneBytesLaunched += sceneBytes;
	}
	if (!okLaunch) {
		// nothing of this launch may be trusted: its requests complete now, failed
		recordFailure(std::string("serve: a launch could not be issued: ") + cadrGetLastErrorMessage());
		(void)hipGetLastError();
		{
			std::lock_guard<std::mutex> lock(m);
			complete(s, CADR_ERROR_DEVICE, nowNs());
		}
		cvDone.notify_all();
		return true;
	}
	s.busy = true;
	return true;
}

void cadrServer::run()
{
	(void)hipSetDevice(device);
	while (!stop) {
		bool did = retire();
		for (Slot& s : slots) {
			if (!s.busy && launch(s)) did = true;
		}
		if (!did) {
			bool idle = true;
			for (Slot& s : slots) idle = idle && !s.busy;
			if (idle) {
				// nothing running: sleep until a submission or a release could let a launch through
				std::unique_lock<std::mutex> lock(m);
				cvWork.wait_for(lock, std::chrono::milliseconds(1), [&] { return stop.load() || wake != wakeSeen; });
			} else {
				std::this_thread::yield();
			}
		}
	}
}

int cadrServer::submitAt(uint32_t scene, const cadrView* view, uint64_t arrivalNs, uint32_t tag, uint64_t* id)
{
	CADR_REQUIRE(view && id, "cadrServerSubmit: null argument");
	CADR_REQUIRE(scene < scenes.size(), "cadrServerSubmit: unknown scene");
	if (scenes[scene].bytesPerView() > pool.size())
		return setError(CADR_ERROR_OUT_OF_RESOURCES, "cadrServerSubmit: the request's output exceeds the budget");
	auto r = std::make_unique<Request>();
	r->scene = scene;
	r->view = *view;
	r->submitNs = arrivalNs;
	r->tag = tag;
	{
		std::lock_guard<std::mutex> lock(m);
		if (queue.size() >= cfg.maxQueuedRequests)
			return setError(CADR_ERROR_OUT_OF_RESOURCES, "cadrServerSubmit: the request queue is full");
		r->id = nextId++;
		*id = r->id;
		queue.push_back(r.get());
		inFlight++;
		wake++;
		requests[r->id] = std::move(r);
	}
	cvWork.notify_one();
	return CADR_SUCCESS;
}

CADR_API int cadrServerCreate(const cadrServerConfig* config, cadrServer** server)
{
	CADR_REQUIRE(config && server, "cadrServerCreate: null argument");
	*server = nullptr;
	const uint32_t tile = cadrCullTileSize(&config->cull);
	CADR_REQUIRE(tile > 0, "cadrServerCreate: unsupported cull config");
	CADR_REQUIRE(config->maxBatchViews > 0 && config->maxInflightBatches > 0 && config->outputBudgetBytes > 0,
	             "cadrServerCreate: maxBatchViews, maxInflightBatches and outputBudgetBytes must be > 0");
	auto s = std::make_unique<cadrServer>();
	s->cfg = *config;
	if (s->cfg.maxQueuedRequests == 0) s->cfg.maxQueuedRequests = 1u << 20;
	s->tileSize = tile;
	CADR_CHECK_HIP(hipGetDevice(&s->device), "hipGetDevice");
	const int rc = s->pool.init(config->outputBudgetBytes);
	if (rc != CADR_SUCCESS) return rc;
	*server = s.release();
	return CADR_SUCCESS;
}

CADR_API int cadrServerDestroy(cadrServer* server)
{
	delete server;
	return CADR_SUCCESS;
}

CADR_API int cadrServerAddScene(cadrServer* s, const cadrServerScene* scene, uint32_t* sceneId)
{
	CADR_REQUIRE(s && scene && sceneId, "cadrServerAddScene: null argument");
	CADR_REQUIRE(!s->started, "cadrServerAddScene: the server is already started");
	CADR_REQUIRE(scene->numStateSets == 0 || scene->stateSetCounts, "cadrServerAddScene: stateSetCounts is NULL");
	uint64_t total = 0;
	for (uint32_t k = 0; k < scene->numStateSets; k++) total += scene->stateSetCounts[k];
	CADR_REQUIRE(total == scene->numDrawables, "cadrServerAddScene: StateSet counts do not sum to numDrawables");
	CADR_REQUIRE(scene->spheres % 16 == 0 && scene->meshes % 16 == 0 && scene->meshIndex % 4 == 0,
	             "cadrServerAddScene: misaligned scene arrays");
	SceneRec rec;
	rec.desc = *scene;
	rec.counts.assign(scene->stateSetCounts, scene->stateSetCounts + scene->numStateSets);
	rec.desc.stateSetCounts = nullptr;
	uint32_t n = 0;
	int rc = cadrBuildCullTiles(rec.counts.data(), scene->numStateSets, s->tileSize, nullptr, 0, &n);
	if (rc != CADR_SUCCESS) return rc;
	std::vector<cadrCullTile> tiles(n);
	rc = cadrBuildCullTiles(rec.counts.data(), scene->numStateSets, s->tileSize, tiles.data(), n, &n);
	if (rc != CADR_SUCCESS) return rc;
	void* d = nullptr;
	CADR_CHECK_HIP(hipMalloc(&d, std::max<size_t>(n, 1) * sizeof(cadrCullTile)), "hipMalloc(tiles)");
	rec.tiles = reinterpret_cast<uint64_t>(d);
	CADR_CHECK_HIP(hipMemcpy(d, tiles.data(), n * sizeof(cadrCullTile), hipMemcpyHostToDevice), "hipMemcpy(tiles)");
	rec.numTiles = n;
	s->maxStateSets = std::max(s->maxStateSets, scene->numStateSets);
	s->maxTiles = std::max(s->maxTiles, n);
	*sceneId = uint32_t(s->scenes.size());
	s->scenes.push_back(std::move(rec));
	return CADR_SUCCESS;
}

CADR_API int cadrServerStart(cadrServer* s)
{
	CADR_REQUIRE(s, "cadrServerStart: server is NULL");
	CADR_REQUIRE(!s->started, "cadrServerStart: already started");
	CADR_REQUIRE(!s->scenes.empty(), "cadrServerStart: no scenes");
	const uint32_t V = s->cfg.maxBatchViews;
	s->slots.resize(s->cfg.maxInflightBatches);
	for (Slot& sl : s->slots) {
		CADR_CHECK_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "hipStreamCreate");
		CADR_CHECK_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
		const int rc = cadrCullWorkspaceCreate(uint64_t(V) * s->maxTiles, V, &sl.ws);
		if (rc != CADR_SUCCESS) return rc;
		void* p = nullptr;
		CADR_CHECK_HIP(hipMalloc(&p, uint64_t(V) * sizeof(cadrView)), "hipMalloc(views)");
		sl.dViews = reinterpret_cast<uint64_t>(p);
		CADR_CHECK_HIP(hipMalloc(&p, uint64_t(V) * sizeof(cadrCullJob)), "hipMalloc(jobs)");
		sl.dJobs = reinterpret_cast<uint64_t>(p);
		CADR_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hViews), uint64_t(V) * sizeof(cadrView),
		                             hipHostMallocDefault), "hipHostMalloc(views)");
		CADR_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hJobs), uint64_t(V) * sizeof(cadrCullJob),
		                             hipHostMallocDefault), "hipHostMalloc(jobs)");
		CADR_CHECK_HIP(hipHostMalloc(reinterpret_cast<void**>(&sl.hCounts), uint64_t(V) * s->maxStateSets * 4u,
		                             hipHostMallocDefault), "hipHostMalloc(counts)");
	}
	s->started = true;
	s->worker = std::thread([s] { s->run(); });
	return CADR_SUCCESS;
}

CADR_API int cadrServerSubmit(cadrServer* s, uint32_t sceneId, const cadrView* view, uint64_t* requestId)
{
	CADR_REQUIRE(s && s->started, "cadrServerSubmit: the server is not started");
	return s->submitAt(sceneId, view, nowNs(), 0u, requestId);
}

CADR_API int cadrServerPoll(cadrServer* s, cadrServerResult* results, uint32_t capacity, uint32_t* count)
{
	CADR_REQUIRE(s && count && (capacity == 0 || results), "cadrServerPoll: null argument");
	std::lock_guard<std::mutex> lock(s->m);
	uint32_t n = 0;
	while (n < capacity && !s->completed.empty()) {
		const uint64_t id = s->completed.front();
		s->completed.pop_front();
		auto it = s->requests.find(id);
		if (it == s->requests.end()) continue;
		const Request& q = *it->second;
		cadrServerResult& r = results[n++];
		r.requestId = q.id;
		r.sceneId = q.scene;
		r.numStateSets = s->scenes[q.scene].desc.numStateSets;
		r.commands = q.commands;
		r.counts = q.counts;
		r.visible = q.visible;
		r.submitNs = q.submitNs;
		r.completeNs = q.completeNs;
		r.status = q.status;
		r.tag = q.tag;
		if (q.status != CADR_SUCCESS) setError(q.status, s->workerMessage);
	}
	*count = n;
	return CADR_SUCCESS;
}

CADR_API int cadrServerCopyCounts(cadrServer* s, uint64_t requestId, uint32_t* counts, uint32_t capacity)
{
	CADR_REQUIRE(s && counts, "cadrServerCopyCounts: null argument");
	std::lock_guard<std::mutex> lock(s->m);
	auto it = s->requests.find(requestId);
	CADR_REQUIRE(it != s->requests.end() && it->second->done, "cadrServerCopyCounts: no completed request with that id");
	CADR_REQUIRE(capacity >= it->second->hostCounts.size(), "cadrServerCopyCounts: capacity too small");
	std::copy(it->second->hostCounts.begin(), it->second->hostCounts.end(), counts);
	return CADR_SUCCESS;
}

CADR_API int cadrServerRelease(cadrServer* s, uint64_t requestId)
{
	CADR_REQUIRE(s, "cadrServerRelease: server is NULL");
	{
		std::lock_guard<std::mutex> lock(s->m);
		auto it = s->requests.find(requestId);
		CADR_REQUIRE(it != s->requests.end() && it->second->done, "cadrServerRelease: no completed request with that id");
		OutputBlock& b = s->blocks[it->second->block];
		if (--b.refs == 0) {
			s->pool.release(b.addr, b.bytes);
			s->freeBlocks.push_back(it->second->block);
		}
		s->requests.erase(it);
		s->wake++;
	}
	s->cvWork.notify_one();  // memory freed: queued requests may fit now
	return CADR_SUCCESS;
}

CADR_API int cadrServerDrain(cadrServer* s, uint32_t timeoutMs)
{
	CADR_REQUIRE(s, "cadrServerDrain: server is NULL");
	std::unique_lock<std::mutex> lock(s->m);
	const bool ok = s->cvDone.wait_for(lock, std::chrono::milliseconds(timeoutMs), [&] { return s->inFlight == 0; });
	if (!ok) return setError(CADR_ERROR_TIMEOUT, "cadrServerDrain: requests still running");
	return CADR_SUCCESS;
}

// ------------------------------------------------------------------ load

namespace {

uint64_t splitmix(uint64_t& x)
{
	uint64_t z = (x += 0x9E3779B97F4A7C15ull);
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

double unit(uint64_t& x) { return double(splitmix(x) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

CADR_API int cadrServerRunLoad(cadrServer* s, const cadrLoadSpec* spec, cadrLoadStats* stats)
{
	CADR_REQUIRE(s && s->started && spec && stats, "cadrServerRunLoad: null argument or server not started");
	CADR_REQUIRE(spec->views && spec->numViews > 0, "cadrServerRunLoad: no views");
	CADR_REQUIRE(spec->requestsPerSecond > 0 && spec->durationSeconds > 0, "cadrServerRunLoad: rate and duration must be > 0");
	CADR_REQUIRE(spec->sampleEvery == 0 || spec->sampleCapacity == 0 || spec->samples,
	             "cadrServerRunLoad: samples is NULL");
	const uint32_t S = uint32_t(s->scenes.size());
	const uint32_t hot = std::min(spec->numHotScenes, S);
	*stats = cadrLoadStats{};
	const uint64_t launches0 = s->launches, views0 = s->launchedViews, sceneBytes0 = s->sceneBytesLaunched;

	// arrival schedule (Poisson process) and request contents
	uint64_t rng = spec->seed;
	struct Arrival {
		uint64_t t;
		uint32_t scene, view;
	};
	std::vector<Arrival> arrivals;
	double t = 0.0;
	for (;;) {
		t += -std::log(1.0 - unit(rng)) / spec->requestsPerSecond;
		if (t >= spec->durationSeconds) break;
		uint32_t scene;
		if (hot > 0 && (hot == S || unit(rng) < spec->hotFraction))
			scene = uint32_t(unit(rng) * hot);
		else
			scene = hot + uint32_t(unit(rng) * (S - hot));
		scene = std::min(scene, S - 1);
		const uint32_t view = std::min(uint32_t(unit(rng) * spec->numViews), spec->numViews - 1);
		arrivals.push_back({uint64_t(t * 1e9), scene, view});
	}

	std::atomic<uint64_t> submitted{0}, refused{0};
	std::atomic<bool> producing{true};
	std::vector<double> latMs;
	latMs.reserve(arrivals.size());
	uint64_t visible = 0, tests = 0, lastComplete = 0;
	const uint64_t start = nowNs() + 2000000;  // first arrival 2 ms from now

	std::atomic<bool> stalled{false}, failed{false};
	uint64_t sampled = 0, failedRequests = 0;
	std::string failure;
	std::thread consumer([&] {
		std::vector<cadrServerResult> res(1024);
		uint64_t done = 0, lastProgress = nowNs();
		for (;;) {
			uint32_t n = 0;
			if (cadrServerPoll(s, res.data(), uint32_t(res.size()), &n) != CADR_SUCCESS) {
				failed = true;
				break;
			}
			if (n) lastProgress = nowNs();
			if (!producing && nowNs() - lastProgress > 30000000000ull) {  // nothing completed for 30 s
				stalled = true;
				break;
			}
			for (uint32_t i = 0; i < n; i++) {
				if (res[i].status != CADR_SUCCESS) {
					if (failedRequests++ == 0) failure = cadrGetLastErrorMessage();
				}
				latMs.push_back(double(res[i].completeNs - res[i].submitNs) * 1e-6);
				visible += res[i].visible;
				tests += s->scenes[res[i].sceneId].desc.numDrawables;
				lastComplete = std::max(lastComplete, res[i].completeNs);
				const uint64_t k = done + i;
				if (spec->sampleEvery && k % spec->sampleEvery == 0 && sampled < spec->sampleCapacity)
					spec->samples[sampled++].result = res[i];  // kept: the caller releases it
				else
					cadrServerRelease(s, res[i].requestId);
			}
			done += n;
			if (!producing && done >= submitted.load()) break;
			if (n == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
		}
	});

	for (const Arrival& a : arrivals) {
		const uint64_t due = start + a.t;
		for (;;) {
			const uint64_t now = nowNs();
			if (now >= due) break;
			if (due - now > 200000) std::this_thread::sleep_for(std::chrono::microseconds((due - now) / 2000));
		}
		uint64_t id = 0;
		// latency counts from the scheduled arrival, so a late submission is not hidden
		if (s->submitAt(a.scene, &spec->views[a.view], due, a.view, &id) == CADR_SUCCESS)
			submitted++;
		else
			refused++;
	}
	producing = false;
	consumer.join();
	if (failed) return setError(CADR_ERROR_DEVICE, "cadrServerRunLoad: polling failed");
	if (stalled) return setError(CADR_ERROR_TIMEOUT, "cadrServerRunLoad: requests stopped completing");
	stats->sampled = sampled;
	stats->failed = failedRequests;

	stats->submitted = submitted;
	stats->refused = refused;
	stats->completed = latMs.size();
	stats->visible = visible;
	stats->launches = s->launches - launches0;
	stats->sceneBytesRead = s->sceneBytesLaunched - sceneBytes0;
	stats->meanViewsPerLaunch = stats->launches ? double(s->launchedViews - views0) / double(stats->launches) : 0.0;
	if (!latMs.empty()) {
		std::sort(latMs.begin(), latMs.end());
		auto pct = [&](double p) { return latMs[std::min(latMs.size() - 1, size_t(p * double(latMs.size())))]; };
		stats->p50Ms = pct(0.50);
		stats->p99Ms = pct(0.99);
		stats->p999Ms = pct(0.999);
		stats->maxMs = latMs.back();
		double sum = 0;
		for (double x : latMs) sum += x;
		stats->meanMs = sum / double(latMs.size());
		stats->elapsedSeconds = double(lastComplete - (start + arrivals.front().t)) * 1e-9;
		stats->requestsPerSecond = double(latMs.size()) / stats->elapsedSeconds;
		stats->drawablesPerSecond = double(tests) / stats->elapsedSeconds;
	}
	if (failedRequests)
		return setError(CADR_ERROR_DEVICE, "cadrServerRunLoad: " + std::to_string(failedRequests) +
		                                       " requests failed on the device: " + failure);
	return CADR_SUCCESS;
}
