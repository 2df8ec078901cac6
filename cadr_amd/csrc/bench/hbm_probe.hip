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

int main()
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
