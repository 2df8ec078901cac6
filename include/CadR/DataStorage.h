// GPU data storage of the CadR facade on HIP memory: handle table, sub-allocated
// data blocks and staged uploads.
//
// Mirrors the reference's DataStorage / DataAllocation / HandlelessAllocation
// / HandleTable (src/CadR/DataStorage.h, DataAllocation.h, HandleTable.h):
//   - a DataAllocation owns a handle from the HandleTable for its whole life;
//     alloc() gives it new memory (a new device address under the same handle,
//     DataAllocation.cpp:18-50) and returns host staging memory to fill;
//   - handles number from 1 (0 is the null handle), and the table deepens from
//     1 to 2 to 3 levels of 2048-entry tables when handle 2048 and handle
//     4,194,304 are created (HandleTable.cpp:113-338), so every handle resolves
//     on the device exactly as processDrawables.comp's lookupHandle expects;
//   - allocations of 64 B or more are 64-byte aligned, smaller ones 16-byte
//     aligned (CircularAllocationMemory.h:565-568);
//   - staged bytes reach the device in executeCopyOperations().
// Device memory comes from a MemoryProvider: HIP device memory for the
// product, host memory for the device-less Renderer (bookkeeping only, as
// the reference's Renderer() constructor, Renderer.cpp:64-74).
#pragma once

#include "Export.h"

#include <array>
#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <set>
#include <utility>
#include <vector>

namespace CadR {

class Renderer;

struct CADR_EXPORT MemoryProvider {
	virtual ~MemoryProvider() = default;
	virtual uint64_t allocate(size_t bytes) = 0;   // throws OutOfResources
	virtual void release(uint64_t address) noexcept = 0;
	// copies that complete before the call returns
	virtual void upload(uint64_t dst, const void* src, size_t bytes) = 0;
	virtual void download(void* dst, uint64_t src, size_t bytes) = 0;
	virtual bool isDevice() const = 0;
};

CADR_EXPORT MemoryProvider& hipMemoryProvider();
CADR_EXPORT MemoryProvider& hostMemoryProvider();

struct DataAllocationRecord {
	uint64_t deviceAddress = 0;
	size_t size = 0;
	uint32_t block = ~0u;             // owning memory block
	uint32_t padding = 0;             // alignment bytes before deviceAddress that belong to it
	size_t stagingFrameNumber = size_t(-2);
	uint8_t* stagingData = nullptr;   // host bytes to upload, valid until executeCopyOperations()
	size_t pendingIndex = 0;          // position in the storage's upload list while stagingData is set
};

class CADR_EXPORT HandleTable {
public:
	static constexpr unsigned numHandlesPerTable = 2048;
	static constexpr unsigned handleBitsLevelShift = 11;
	static constexpr unsigned handleBitsLevelMask = 0x07ff;

	explicit HandleTable(class DataStorage& storage) noexcept : _storage(&storage) {}
	~HandleTable() noexcept;
	uint64_t create();
	void destroy(uint64_t) noexcept {}   // handles are not reused (HandleTable.h:91)
	void set(uint64_t handle, uint64_t address);
	uint64_t get(uint64_t handle) const;  // host mirror lookup (0 for unknown handles)
	unsigned handleLevel() const { return _level; }
	uint64_t rootTableDeviceAddress() const;
	uint64_t highestHandle() const { return _highest; }
	size_t numTables() const { return _tables.size(); }
	// uploads every table changed since the last call
	void uploadDirty();

private:
	struct Table {
		std::array<uint64_t, numHandlesPerTable> entries{};
		DataAllocationRecord* memory = nullptr;
		bool dirty = true;
	};
	Table* newTable();
	Table* tableOf(uint64_t handle) const;
	void setEntry(Table* t, unsigned index, uint64_t value);

	class DataStorage* _storage;
	std::vector<std::unique_ptr<Table>> _tables;
	Table* _root = nullptr;                 // level-1: the only leaf table; level 2/3: routing table
	std::vector<Table*> _l1;                // level 3: second-level routing tables by top index
	std::vector<Table*> _leafByIndex;       // leaf table of handle >> 11
	uint64_t _highest = 0;
	unsigned _level = 0;
};

class CADR_EXPORT DataStorage {
public:
	explicit DataStorage(Renderer& r, MemoryProvider& provider);
	~DataStorage() noexcept;
	DataStorage(const DataStorage&) = delete;
	DataStorage& operator=(const DataStorage&) = delete;

	Renderer& renderer() const { return *_renderer; }
	MemoryProvider& provider() const { return *_provider; }

	// allocation (DataStorage::alloc / realloc, DataStorage.cpp:111-165)
	DataAllocationRecord* alloc(size_t numBytes);
	DataAllocationRecord* realloc(DataAllocationRecord* record, size_t numBytes);
	void free(DataAllocationRecord* record) noexcept;
	DataAllocationRecord* zeroSizeAllocationRecord() noexcept { return &_zeroRecord; }

	// staging: host bytes that reach the device at executeCopyOperations()
	uint8_t* stage(DataAllocationRecord* record);
	size_t recordUploads();   // performs pending uploads; returns the number of bytes

	// handles
	uint64_t createHandle() { return _handleTable.create(); }
	void destroyHandle(uint64_t h) noexcept { _handleTable.destroy(h); }
	void setHandle(uint64_t h, uint64_t address) { _handleTable.set(h, address); }
	unsigned handleLevel() const { return _handleTable.handleLevel(); }
	uint64_t handleTableDeviceAddress() const { return _handleTable.rootTableDeviceAddress(); }
	const HandleTable& handleTable() const { return _handleTable; }

	size_t numBlocks() const { return _liveBlocks; }
	size_t liveAllocations() const { return _live; }

	static constexpr size_t smallMemorySize = 64 << 10;   // Renderer.h:104-106
	static constexpr size_t mediumMemorySize = 2 << 20;
	static constexpr size_t largeMemorySize = 32 << 20;

private:
	// Free space of a block: the space past its last allocation (`tail`) and
	// coalesced holes before it, indexed by offset and by size.  An
	// allocation takes the smallest hole it fits in (at its aligned start),
	// else the tail, so an empty block is filled front to back with the alignment
	// ladder of CircularAllocationMemory.h:565-568, as the reference places
	// consecutive allocations, and space freed anywhere in the block is used
	// again (the reference's circular buffer reuses only what lies behind its
	// oldest live allocation, CircularAllocationMemory.h:522-560).
	struct Block {
		uint64_t base = 0;   // 0: released
		size_t size = 0;
		size_t live = 0;
		size_t tail = 0;     // [tail, size) is free; the ranges below hold the holes before it
		std::map<size_t, size_t> freeByOffset;           // offset -> size
		std::set<std::pair<size_t, size_t>> freeBySize;  // (size, offset)
	};
	// aligned offset of a new allocation in block b, or false
	static bool fit(const Block& b, size_t numBytes, size_t align, size_t& offset);
	// takes [at, at + numBytes); returns the alignment padding before `at`
	// that goes with it (free space too small to matter)
	static size_t takeRange(Block& b, size_t at, size_t numBytes);
	static void returnRange(Block& b, size_t at, size_t numBytes);
	uint32_t blockFor(size_t numBytes, size_t align, size_t& offset);
	uint32_t newBlock(size_t size);
	uint32_t _firstBlock = ~0u, _secondBlock = ~0u;  // allocation order of DataStorage.cpp:31-101
	size_t _liveBlocks = 0;

	Renderer* _renderer;
	MemoryProvider* _provider;
	std::vector<Block> _blocks;
	std::vector<std::unique_ptr<DataAllocationRecord>> _recordPool;
	std::vector<DataAllocationRecord*> _freeRecords;
	std::vector<DataAllocationRecord*> _pending;   // records with staged bytes
	std::vector<std::unique_ptr<uint8_t[]>> _stagingChunks;
	DataAllocationRecord _zeroRecord;
	HandleTable _handleTable;
	size_t _live = 0;
	friend class HandleTable;
};

// A piece of device memory with a handle (DataAllocation.h:40-120).
class CADR_EXPORT DataAllocation {
public:
	explicit DataAllocation(DataStorage& storage);
	DataAllocation(DataAllocation&& other) noexcept;
	DataAllocation& operator=(DataAllocation&& rhs) noexcept;
	DataAllocation(const DataAllocation&) = delete;
	DataAllocation& operator=(const DataAllocation&) = delete;
	~DataAllocation() noexcept;

	// new memory under the same handle; returns host staging bytes to fill
	void* alloc(size_t numBytes);
	void free() noexcept;
	void upload(const void* ptr, size_t numBytes);
	template <typename T> T* editNewContent(size_t count) { return static_cast<T*>(alloc(sizeof(T) * count)); }

	uint64_t handle() const { return _handle; }
	uint64_t deviceAddress() const { return _record->deviceAddress; }
	size_t size() const { return _record->size; }
	DataStorage& dataStorage() const { return *_storage; }

private:
	DataStorage* _storage;
	DataAllocationRecord* _record;
	uint64_t _handle;
};

// Device memory without a handle (DataAllocation.h:124-170).
class CADR_EXPORT HandlelessAllocation {
public:
	explicit HandlelessAllocation(DataStorage& storage) noexcept;
	~HandlelessAllocation() noexcept;
	HandlelessAllocation(const HandlelessAllocation&) = delete;
	HandlelessAllocation& operator=(const HandlelessAllocation&) = delete;
	void* alloc(size_t numBytes);
	void free() noexcept;
	uint64_t deviceAddress() const { return _record->deviceAddress; }
	size_t size() const { return _record->size; }

private:
	DataStorage* _storage;
	DataAllocationRecord* _record;
};

}  // namespace CadR
