// DataStorage, HandleTable and DataAllocation of the CadR facade.
#include "../../../include/CadR/DataStorage.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../../include/CadR/Exceptions.h"
#include "../../../include/CadR/Renderer.h"
#include "../../../include/cadr_hip.h"

namespace CadR {

void throwOnError(int code, const char* what)
{
	if (code == CADR_SUCCESS) return;
	std::string msg = std::string(what) + ": " + cadrGetLastErrorMessage();
	switch (code) {
	case CADR_ERROR_INVALID_VALUE: throw LogicError(msg);
	case CADR_ERROR_OUT_OF_RESOURCES: throw OutOfResources(msg);
	case CADR_ERROR_TIMEOUT: throw Timeout(msg);
	default: throw DeviceError(msg);
	}
}

// ------------------------------------------------------------ providers

namespace {

struct HipProvider final : MemoryProvider {
	uint64_t allocate(size_t bytes) override
	{
		void* p = nullptr;
		if (hipMalloc(&p, bytes) != hipSuccess) {
			(void)hipGetLastError();
			throw OutOfResources("CadR::DataStorage: cannot allocate " + std::to_string(bytes) + " bytes of HIP memory");
		}
		return reinterpret_cast<uint64_t>(p);
	}
	void release(uint64_t a) noexcept override { (void)hipFree(reinterpret_cast<void*>(a)); }
	void upload(uint64_t dst, const void* src, size_t bytes) override
	{
		if (hipMemcpy(reinterpret_cast<void*>(dst), src, bytes, hipMemcpyHostToDevice) != hipSuccess)
			throw DeviceError(std::string("hipMemcpy(HtoD): ") + hipGetErrorString(hipGetLastError()));
	}
	void download(void* dst, uint64_t src, size_t bytes) override
	{
		if (hipMemcpy(dst, reinterpret_cast<const void*>(src), bytes, hipMemcpyDeviceToHost) != hipSuccess)
			throw DeviceError(std::string("hipMemcpy(DtoH): ") + hipGetErrorString(hipGetLastError()));
	}
	bool isDevice() const override { return true; }
};

struct HostProvider final : MemoryProvider {
	uint64_t allocate(size_t bytes) override
	{
		void* p = std::aligned_alloc(256, (bytes + 255) / 256 * 256);
		if (!p) throw OutOfResources("CadR::DataStorage: cannot allocate host memory");
		std::memset(p, 0, bytes);
		return reinterpret_cast<uint64_t>(p);
	}
	void release(uint64_t a) noexcept override { std::free(reinterpret_cast<void*>(a)); }
	void upload(uint64_t dst, const void* src, size_t bytes) override
	{
		std::memcpy(reinterpret_cast<void*>(dst), src, bytes);
	}
	void download(void* dst, uint64_t src, size_t bytes) override
	{
		std::memcpy(dst, reinterpret_cast<const void*>(src), bytes);
	}
	bool isDevice() const override { return false; }
};

}  // namespace

MemoryProvider& hipMemoryProvider()
{
	static HipProvider p;
	return p;
}

MemoryProvider& hostMemoryProvider()
{
	static HostProvider p;
	return p;
}

// ---------------------------------------------------------- DataStorage

DataStorage::DataStorage(Renderer& r, MemoryProvider& provider)
    : _renderer(&r), _provider(&provider), _handleTable(*this)
{
}

DataStorage::~DataStorage() noexcept
{
	// the handle table (a member, destroyed after this body) frees its tables
	// into blocks that are gone by then
	for (Block& b : _blocks) {
		if (b.base) _provider->release(b.base);
		b.base = 0;
	}
}

// Alignment ladder of the reference's allocator: 64 B for allocations of 64 B
// and more, 16 B otherwise (CircularAllocationMemory.h:565-568).
static size_t alignmentFor(size_t numBytes) { return numBytes >= 64 ? 64 : 16; }

// The smallest hole that holds numBytes at an aligned start, else the tail.
// Holes of exactly numBytes or a little more may not fit once aligned, so a
// few candidates are tried before one of numBytes + align - 1 is taken, which
// always fits.
bool DataStorage::fit(const Block& b, size_t numBytes, size_t align, size_t& offset)
{
	if (!b.freeBySize.empty()) {
		auto it = b.freeBySize.lower_bound({numBytes, 0});
		for (int tries = 0; it != b.freeBySize.end() && tries < 8; ++it, ++tries) {
			const size_t at = (it->second + align - 1) / align * align;
			if (at + numBytes <= it->second + it->first) {
				offset = at;
				return true;
			}
		}
		it = b.freeBySize.lower_bound({numBytes + align - 1, 0});
		if (it != b.freeBySize.end()) {
			offset = (it->second + align - 1) / align * align;
			return true;
		}
	}
	const size_t at = (b.tail + align - 1) / align * align;
	if (at + numBytes > b.size) return false;
	offset = at;
	return true;
}

namespace {

// A hole [off, off + len) in both indexes, reusing the nodes of `byOffset`
// and `bySize` when given (node handles of a hole just taken out).
template <class M, class S>
void putHole(M& byOffset, S& bySize, size_t off, size_t len, typename M::node_type&& o = {},
             typename S::node_type&& z = {})
{
	if (o) {
		o.key() = off;
		o.mapped() = len;
		byOffset.insert(std::move(o));
	} else {
		byOffset.emplace(off, len);
	}
	if (z) {
		z.value() = {len, off};
		bySize.insert(std::move(z));
	} else {
		bySize.emplace(len, off);
	}
}

}  // namespace

// Removes [at, at + numBytes) from the free space holding it.  Space left
// after it stays free; the alignment padding before it (under 64 bytes)
// goes with the allocation and is returned with it.
size_t DataStorage::takeRange(Block& b, size_t at, size_t numBytes)
{
	if (at >= b.tail) {
		const size_t pad = at - b.tail;
		b.tail = at + numBytes;
		return pad;
	}
	auto r = std::prev(b.freeByOffset.upper_bound(at));
	const size_t start = r->first, len = r->second, end = at + numBytes, rend = start + len;
	auto o = b.freeByOffset.extract(r);
	auto z = b.freeBySize.extract({len, start});
	if (rend > end) putHole(b.freeByOffset, b.freeBySize, end, rend - end, std::move(o), std::move(z));
	return at - start;
}

// Gives [at, at + numBytes) back: to the tail when it ends there, else as a
// hole merged with free neighbours.
void DataStorage::returnRange(Block& b, size_t at, size_t numBytes)
{
	size_t start = at, end = at + numBytes;
	auto next = b.freeByOffset.lower_bound(at);
	decltype(b.freeByOffset.extract(next)) o;
	decltype(b.freeBySize.extract({0, 0})) z;
	if (next != b.freeByOffset.end() && next->first == end) {
		end += next->second;
		z = b.freeBySize.extract({next->second, next->first});
		auto after = std::next(next);
		o = b.freeByOffset.extract(next);
		next = after;
	}
	if (next != b.freeByOffset.begin()) {
		auto prev = std::prev(next);
		if (prev->first + prev->second == start) {
			start = prev->first;
			auto z2 = b.freeBySize.extract({prev->second, prev->first});
			auto o2 = b.freeByOffset.extract(prev);
			if (!o) o = std::move(o2);
			if (!z) z = std::move(z2);
		}
	}
	if (end == b.tail) {
		b.tail = start;  // the space joins the tail; a hole right before it was merged above
		return;
	}
	putHole(b.freeByOffset, b.freeBySize, start, end - start, std::move(o), std::move(z));
}

uint32_t DataStorage::newBlock(size_t size)
{
	Block b;
	b.base = _provider->allocate(size);
	b.size = size;
	_liveBlocks++;
	for (uint32_t i = 0; i < _blocks.size(); i++)
		if (_blocks[i].base == 0) {
			_blocks[i] = std::move(b);
			return i;
		}
	_blocks.push_back(std::move(b));
	return uint32_t(_blocks.size() - 1);
}

// Block choice of DataStorage::allocInternal (DataStorage.cpp:31-101): the
// first allocation block is tried, then the second; the first block is
// 64 KiB (or larger for a large first allocation), the second 2 MiB.  When
// neither has room, the other blocks still holding allocations are tried in
// turn (the reference never returns to them), and only then a new large
// block (32 MiB or the allocation's size) becomes the second and the old
// second the first.
uint32_t DataStorage::blockFor(size_t numBytes, size_t align, size_t& offset)
{
	if (_firstBlock == ~0u)
		_firstBlock = newBlock(numBytes < smallMemorySize    ? smallMemorySize
		                       : numBytes < mediumMemorySize ? mediumMemorySize
		                                                     : std::max(numBytes, largeMemorySize));
	if (fit(_blocks[_firstBlock], numBytes, align, offset)) return _firstBlock;
	if (_secondBlock == ~0u)
		_secondBlock = newBlock(numBytes < mediumMemorySize ? mediumMemorySize : std::max(numBytes, largeMemorySize));
	if (fit(_blocks[_secondBlock], numBytes, align, offset)) return _secondBlock;
	for (uint32_t i = 0; i < _blocks.size(); i++)
		if (_blocks[i].base && i != _firstBlock && i != _secondBlock && fit(_blocks[i], numBytes, align, offset))
			return i;
	const uint32_t b = newBlock(std::max(numBytes, largeMemorySize));
	_firstBlock = _secondBlock;
	_secondBlock = b;
	offset = 0;
	return b;
}

DataAllocationRecord* DataStorage::alloc(size_t numBytes)
{
	if (numBytes == 0) return &_zeroRecord;  // null-object pattern (DataStorage.cpp:125-128)
	const size_t align = alignmentFor(numBytes);
	size_t at = 0;
	const uint32_t bi = blockFor(numBytes, align, at);
	Block& b = _blocks[bi];
	DataAllocationRecord* r;
	if (!_freeRecords.empty()) {
		r = _freeRecords.back();
		_freeRecords.pop_back();
	} else {
		_recordPool.push_back(std::make_unique<DataAllocationRecord>());
		r = _recordPool.back().get();
	}
	const size_t pad = takeRange(b, at, numBytes);
	b.live++;
	*r = DataAllocationRecord{};
	r->deviceAddress = b.base + at;
	r->size = numBytes;
	r->block = bi;
	r->padding = uint32_t(pad);
	r->stagingFrameNumber = _renderer->frameNumber();
	_live++;
	return r;
}

DataAllocationRecord* DataStorage::realloc(DataAllocationRecord* record, size_t numBytes)
{
	if (numBytes == 0) {
		free(record);
		return &_zeroRecord;
	}
	DataAllocationRecord* a = alloc(numBytes);
	free(record);
	return a;
}

void DataStorage::free(DataAllocationRecord* r) noexcept
{
	if (!r || r == &_zeroRecord || r->size == 0) return;
	// drop staged bytes that were never uploaded (the upload list's order
	// does not matter: live allocations never overlap)
	if (r->stagingData) {
		DataAllocationRecord* last = _pending.back();
		_pending[r->pendingIndex] = last;
		last->pendingIndex = r->pendingIndex;
		_pending.pop_back();
	}
	const uint32_t bi = r->block;
	Block& b = _blocks[bi];
	if (b.base) returnRange(b, size_t(r->deviceAddress - b.base) - r->padding, r->size + r->padding);
	_live--;
	*r = DataAllocationRecord{};
	_freeRecords.push_back(r);
	if (--b.live == 0 && bi != _firstBlock && bi != _secondBlock && b.base) {
		// a block outside the allocation order that holds nothing goes back
		_provider->release(b.base);
		b = Block{};
		_liveBlocks--;
	}
}

uint8_t* DataStorage::stage(DataAllocationRecord* r)
{
	if (r->size == 0) return nullptr;
	if (!r->stagingData) {
		_stagingChunks.push_back(std::make_unique<uint8_t[]>(r->size));
		r->stagingData = _stagingChunks.back().get();
		std::memset(r->stagingData, 0, r->size);
		r->pendingIndex = _pending.size();
		_pending.push_back(r);
	}
	return r->stagingData;
}

size_t DataStorage::recordUploads()
{
	size_t bytes = 0;
	_handleTable.uploadDirty();
	for (DataAllocationRecord* r : _pending) {
		if (r->stagingData && r->size) {
			_provider->upload(r->deviceAddress, r->stagingData, r->size);
			bytes += r->size;
		}
		r->stagingData = nullptr;
	}
	_pending.clear();
	_stagingChunks.clear();
	return bytes;
}

// ----------------------------------------------------------- HandleTable
//
// Handles are numbered 1, 2, 3, ... (0 is the null handle).  Level 1: one
// table of 2048 entries indexed by the handle.  Creating handle 2048 adds a
// routing table whose entry i points to the leaf table of handles
// [2048 i, 2048 i + 2047] (HandleTable.cpp:142-200).  Creating handle
// 2048*2048 adds a top routing table over 2048-entry second-level routing
// tables (HandleTable.cpp:203-338).  Entries of tables that do not exist read
// as 0.  Every table is 16 KiB of device memory from the DataStorage.

HandleTable::~HandleTable() noexcept
{
	for (auto& t : _tables)
		if (t->memory) _storage->free(t->memory);
}

HandleTable::Table* HandleTable::newTable()
{
	_tables.push_back(std::make_unique<Table>());
	Table* t = _tables.back().get();
	t->memory = _storage->alloc(numHandlesPerTable * sizeof(uint64_t));
	t->dirty = true;
	return t;
}

void HandleTable::setEntry(Table* t, unsigned index, uint64_t value)
{
	t->entries[index] = value;
	t->dirty = true;
}

uint64_t HandleTable::create()
{
	const uint64_t h = _highest + 1;
	const uint64_t leafIndex = h >> handleBitsLevelShift;
	if (_level == 0) {
		_root = newTable();
		_leafByIndex.push_back(_root);
		_level = 1;
	} else if (_level == 1 && leafIndex == 1) {
		// 2048th handle: routing table over leaf tables 0 and 1
		Table* leaf0 = _root;
		Table* r = newTable();
		Table* leaf1 = newTable();
		setEntry(r, 0, leaf0->memory->deviceAddress);
		setEntry(r, 1, leaf1->memory->deviceAddress);
		_leafByIndex.push_back(leaf1);
		_root = r;
		_level = 2;
	} else if (_level == 2 && (h & handleBitsLevelMask) == 0) {
		if (leafIndex < numHandlesPerTable) {
			Table* leaf = newTable();
			setEntry(_root, unsigned(leafIndex), leaf->memory->deviceAddress);
			_leafByIndex.push_back(leaf);
		} else {
			// 2048*2048th handle: top table over second-level routing tables
			Table* l1old = _root;
			Table* top = newTable();
			Table* l1 = newTable();
			Table* leaf = newTable();
			setEntry(l1, 0, leaf->memory->deviceAddress);
			setEntry(top, 0, l1old->memory->deviceAddress);
			setEntry(top, 1, l1->memory->deviceAddress);
			_l1.push_back(l1old);
			_l1.push_back(l1);
			_leafByIndex.push_back(leaf);
			_root = top;
			_level = 3;
		}
	} else if (_level == 3 && (h & handleBitsLevelMask) == 0) {
		const unsigned topIndex = unsigned(h >> (2 * handleBitsLevelShift));
		const unsigned midIndex = unsigned(leafIndex & handleBitsLevelMask);
		if (topIndex >= numHandlesPerTable)
			throw OutOfResources("CadR::HandleTable: the 3-level table is full (2048^3 - 1 handles)");
		if (midIndex == 0) {
			Table* l1 = newTable();
			setEntry(_root, topIndex, l1->memory->deviceAddress);
			_l1.push_back(l1);
		}
		Table* leaf = newTable();
		setEntry(_l1[topIndex], midIndex, leaf->memory->deviceAddress);
		_leafByIndex.push_back(leaf);
	}
	_highest = h;
	return h;
}

HandleTable::Table* HandleTable::tableOf(uint64_t handle) const
{
	const uint64_t leafIndex = handle >> handleBitsLevelShift;
	if (handle == 0 || handle > _highest || leafIndex >= _leafByIndex.size()) return nullptr;
	return _leafByIndex[leafIndex];
}

void HandleTable::set(uint64_t handle, uint64_t address)
{
	Table* t = tableOf(handle);
	if (!t) throw LogicError("CadR::HandleTable::set(): invalid handle " + std::to_string(handle));
	setEntry(t, unsigned(handle & handleBitsLevelMask), address);
}

uint64_t HandleTable::get(uint64_t handle) const
{
	Table* t = tableOf(handle);
	return t ? t->entries[handle & handleBitsLevelMask] : 0;
}

uint64_t HandleTable::rootTableDeviceAddress() const
{
	return _root ? _root->memory->deviceAddress : 0;
}

void HandleTable::uploadDirty()
{
	for (auto& t : _tables) {
		if (!t->dirty) continue;
		_storage->provider().upload(t->memory->deviceAddress, t->entries.data(), sizeof(t->entries));
		t->dirty = false;
	}
}

// -------------------------------------------------------- DataAllocation

DataAllocation::DataAllocation(DataStorage& storage)
    : _storage(&storage), _record(storage.zeroSizeAllocationRecord()), _handle(storage.createHandle())
{
}

DataAllocation::DataAllocation(DataAllocation&& o) noexcept : _storage(o._storage), _record(o._record), _handle(o._handle)
{
	o._record = o._storage->zeroSizeAllocationRecord();
	o._handle = 0;
}

DataAllocation& DataAllocation::operator=(DataAllocation&& rhs) noexcept
{
	if (this != &rhs) {
		free();
		if (_handle) _storage->destroyHandle(_handle);
		_storage = rhs._storage;
		_record = rhs._record;
		_handle = rhs._handle;
		rhs._record = rhs._storage->zeroSizeAllocationRecord();
		rhs._handle = 0;
	}
	return *this;
}

DataAllocation::~DataAllocation() noexcept
{
	free();
	if (_handle) _storage->destroyHandle(_handle);
}

void* DataAllocation::alloc(size_t numBytes)
{
	// reuse memory allocated earlier in this frame (DataAllocation.cpp:23-27)
	if (_record->size != 0 && _record->stagingFrameNumber == _storage->renderer().frameNumber() &&
	    numBytes <= _record->size && _record->stagingData) {
		return _record->stagingData;
	}
	_record = _storage->realloc(_record, numBytes);
	if (_handle) _storage->setHandle(_handle, _record->deviceAddress);
	return _storage->stage(_record);
}

void DataAllocation::free() noexcept
{
	if (_record->size == 0) return;
	_storage->free(_record);
	_record = _storage->zeroSizeAllocationRecord();
	if (_handle) {
		try {
			_storage->setHandle(_handle, 0);
		} catch (...) {
		}
	}
}

void DataAllocation::upload(const void* ptr, size_t numBytes)
{
	void* p = alloc(numBytes);
	if (numBytes) std::memcpy(p, ptr, numBytes);
}

HandlelessAllocation::HandlelessAllocation(DataStorage& storage) noexcept
    : _storage(&storage), _record(storage.zeroSizeAllocationRecord())
{
}

HandlelessAllocation::~HandlelessAllocation() noexcept { free(); }

void* HandlelessAllocation::alloc(size_t numBytes)
{
	_record = _storage->realloc(_record, numBytes);
	return _storage->stage(_record);
}

void HandlelessAllocation::free() noexcept
{
	if (_record->size == 0) return;
	_storage->free(_record);
	_record = _storage->zeroSizeAllocationRecord();
}

}  // namespace CadR
