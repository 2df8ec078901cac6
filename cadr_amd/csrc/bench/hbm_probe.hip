// Achievable HBM bandwidth of the GPU the benchmark runs on: a streaming
// copy, a read-only pass and a write-only pass over 1-GiB buffers (one float4
// per thread, the fastest of the simple forms measured on MI355X), 20 timed
// launches each after 2 warm-up launches, HIP events.  Prints one JSON line
// with GB/s (bytes moved: copy = read + write).  The kernels' rooflines are
// quoted against the 8 TB/s published peak; this gives the measured ceiling
// beside it.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#define CHECK(x)                                                                        \
	do {                                                                                \
		hipError_t e_ = (x);                                                            \
		if (e_ != hipSuccess) {                                                         \
			fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));              \
			return 1;                                                                   \
		}                                                                               \
	} while (0)

__global__ __launch_bounds__(256) void copyKernel(const float4* __restrict__ a, float4* __restrict__ b, uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i < n) b[i] = a[i];
}

__global__ __launch_bounds__(256) void readKernel(const float4* __restrict__ a, float* __restrict__ sink, uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i < n) {
		const float4 v = a[i];
		if (v.x + v.y + v.z + v.w == 1.0f) sink[0] = v.x;  // never true for the zero-filled input
	}
}

__global__ __launch_bounds__(256) void writeKernel(float4* __restrict__ b, uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i < n) b[i] = make_float4(1.f, 2.f, 3.f, 4.f);
}

// Access patterns of known byte counts for calibrating the L2-to-memory
// counters (rocprofv3 FETCH_SIZE / WRITE_SIZE) beyond the streaming copy:
// every pattern moves each byte of a 1-GiB buffer exactly once, in records of
// RECORD bytes (RECORD / 16 consecutive lanes per record, one float4 each)
// taken in the order of the bijection r -> r * 0x9E3779B1 mod R over the
// R = 2^30 / RECORD records (the multiplier is odd and R a power of two), so
// records that share a 128-byte line are read far apart in time.
__device__ __forceinline__ uint64_t permutedRecord(uint64_t i, uint32_t lanesPerRecord, uint32_t records)
{
	const uint32_t r = uint32_t(i / lanesPerRecord);
	return uint64_t((r * 0x9E3779B1u) & (records - 1u)) * lanesPerRecord + i % lanesPerRecord;
}

template <int RECORD>
__global__ __launch_bounds__(256) void gatherReadKernel(const float4* __restrict__ a, float* __restrict__ sink,
                                                        uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i < n) {
		const float4 v = a[permutedRecord(i, RECORD / 16, uint32_t(n / (RECORD / 16)))];
		if (v.x + v.y + v.z + v.w == 1.0f) sink[0] = v.x;
	}
}

// The first 64 bytes of every 128-byte line, lines in permuted order: half of
// the buffer's bytes, in every line.
__global__ __launch_bounds__(256) void halfLineReadKernel(const float4* __restrict__ a, float* __restrict__ sink,
                                                          uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i < n / 2) {
		const uint64_t j = permutedRecord(i, 4, uint32_t(n / 8));  // 64-B record index space of the lines
		const uint64_t line = j / 4, part = j % 4;
		const float4 v = a[line * 8 + part];
		if (v.x + v.y + v.z + v.w == 1.0f) sink[0] = v.x;
	}
}

template <int RECORD>
__global__ __launch_bounds__(256) void scatterWriteKernel(float4* __restrict__ b, uint64_t n)
{
	const uint64_t i = uint64_t(blockIdx.x) * 256u + threadIdx.x;
	if (i < n) b[permutedRecord(i, RECORD / 16, uint32_t(n / (RECORD / 16)))] = make_float4(1.f, 2.f, 3.f, 4.f);
}

// --fetch-calibration: each pattern twice (the counters are read per
// dispatch), printing its bytes and bandwidth.
static int fetchCalibration(float4* a, float4* b, float* sink, uint64_t bytes)
{
	const uint64_t n = bytes / sizeof(float4);
	const uint32_t blocks = uint32_t((n + 255) / 256);
	hipEvent_t e0, e1;
	CHECK(hipEventCreate(&e0));
	CHECK(hipEventCreate(&e1));
	struct Pattern {
		const char* name;
		uint64_t bytesRead, bytesWritten;
		void (*launch)(float4*, float4*, float*, uint64_t, uint32_t);
	};
	const Pattern patterns[] = {
		{"stream_read", bytes, 0,
		 [](float4* a, float4*, float* s, uint64_t n, uint32_t g) { hipLaunchKernelGGL(readKernel, dim3(g), dim3(256), 0, 0, a, s, n); }},
		{"gather_read_16B", bytes, 0,
		 [](float4* a, float4*, float* s, uint64_t n, uint32_t g) { hipLaunchKernelGGL(gatherReadKernel<16>, dim3(g), dim3(256), 0, 0, a, s, n); }},
		{"gather_read_64B", bytes, 0,
		 [](float4* a, float4*, float* s, uint64_t n, uint32_t g) { hipLaunchKernelGGL(gatherReadKernel<64>, dim3(g), dim3(256), 0, 0, a, s, n); }},
		{"gather_read_128B", bytes, 0,
		 [](float4* a, float4*, float* s, uint64_t n, uint32_t g) { hipLaunchKernelGGL(gatherReadKernel<128>, dim3(g), dim3(256), 0, 0, a, s, n); }},
		{"half_line_read_64B", bytes / 2, 0,
		 [](float4* a, float4*, float* s, uint64_t n, uint32_t g) { hipLaunchKernelGGL(halfLineReadKernel, dim3((g + 1) / 2), dim3(256), 0, 0, a, s, n); }},
		{"stream_write", 0, bytes,
		 [](float4*, float4* b, float*, uint64_t n, uint32_t g) { hipLaunchKernelGGL(writeKernel, dim3(g), dim3(256), 0, 0, b, n); }},
		{"scatter_write_16B", 0, bytes,
		 [](float4*, float4* b, float*, uint64_t n, uint32_t g) { hipLaunchKernelGGL(scatterWriteKernel<16>, dim3(g), dim3(256), 0, 0, b, n); }},
		{"scatter_write_64B", 0, bytes,
		 [](float4*, float4* b, float*, uint64_t n, uint32_t g) { hipLaunchKernelGGL(scatterWriteKernel<64>, dim3(g), dim3(256), 0, 0, b, n); }},
	};
	for (const Pattern& p : patterns) {
		p.launch(a, b, sink, n, blocks);  // warm-up
		CHECK(hipGetLastError());
		CHECK(hipDeviceSynchronize());
		CHECK(hipEventRecord(e0));
		p.launch(a, b, sink, n, blocks);
		CHECK(hipEventRecord(e1));
		CHECK(hipEventSynchronize(e1));
		float ms = 0.f;
		CHECK(hipEventElapsedTime(&ms, e0, e1));
		const double moved = double(p.bytesRead + p.bytesWritten);
		printf("{\"pattern\": \"%s\", \"bytes_read\": %llu, \"bytes_written\": %llu, \"ms\": %.4f, \"gbs\": %.1f}\n",
		       p.name, (unsigned long long)p.bytesRead, (unsigned long long)p.bytesWritten, ms,
		       moved / (double(ms) * 1e-3) / 1e9);
		fflush(stdout);
	}
	CHECK(hipEventDestroy(e0));
	CHECK(hipEventDestroy(e1));
	return 0;
}

int main(int argc, char** argv)
{
	const uint64_t bytes = 1ull << 30, n = bytes / sizeof(float4);
	const uint32_t blocks = uint32_t((n + 255) / 256);
	float4 *a = nullptr, *b = nullptr;
	float* sink = nullptr;
	CHECK(hipMalloc(&a, bytes));
	CHECK(hipMalloc(&b, bytes));
	CHECK(hipMalloc(&sink, 64));
	CHECK(hipMemset(a, 0, bytes));
	CHECK(hipMemset(b, 0, bytes));
	if (argc > 1 && std::strcmp(argv[1], "--fetch-calibration") == 0) {
		const int rc = fetchCalibration(a, b, sink, bytes);
		CHECK(hipFree(a));
		CHECK(hipFree(b));
		CHECK(hipFree(sink));
		return rc;
	}
	hipEvent_t e0, e1;
	CHECK(hipEventCreate(&e0));
	CHECK(hipEventCreate(&e1));
	double gbs[3];
	for (int k = 0; k < 3; k++) {
		auto launch = [&] {
			if (k == 0) hipLaunchKernelGGL(copyKernel, dim3(blocks), dim3(256), 0, 0, a, b, n);
			else if (k == 1) hipLaunchKernelGGL(readKernel, dim3(blocks), dim3(256), 0, 0, a, sink, n);
			else hipLaunchKernelGGL(writeKernel, dim3(blocks), dim3(256), 0, 0, b, n);
		};
		launch();
		launch();
		CHECK(hipGetLastError());
		CHECK(hipDeviceSynchronize());
		const int reps = 20;
		CHECK(hipEventRecord(e0));
		for (int r = 0; r < reps; r++) launch();
		CHECK(hipEventRecord(e1));
		CHECK(hipEventSynchronize(e1));
		float ms = 0.f;
		CHECK(hipEventElapsedTime(&ms, e0, e1));
		const double moved = k == 0 ? 2.0 * bytes : double(bytes);
		gbs[k] = moved / (double(ms) / reps * 1e-3) / 1e9;
	}
	printf("{\"bytes\": %llu, \"copy_gbs\": %.1f, \"read_gbs\": %.1f, \"write_gbs\": %.1f}\n",
	       (unsigned long long)bytes, gbs[0], gbs[1], gbs[2]);
	CHECK(hipFree(a));
	CHECK(hipFree(b));
	CHECK(hipFree(sink));
	return 0;
}
