/*
 * infoli_calib.hip -- measurement helpers for the benchmark write-up.
 *
 *   infoli_calib.x copy   [MiB]      plain device copy of a known size: the
 *                                    reference point for scaling rocprofv3's
 *                                    FETCH_SIZE / WRITE_SIZE counters
 *   infoli_calib.x gather [table] [xcd]
 *                                    random 4-byte gathers from a table of
 *                                    `table` floats (global memory) and, when
 *                                    it fits, from LDS; with `xcd`, each
 *                                    workgroup reads one eighth of the table,
 *                                    chosen by its XCD
 *   infoli_calib.x exp               throughput of the hardware exp2 (v_exp_f32)
 *   infoli_calib.x launch            per-kernel cost inside a hipGraph, in the
 *                                    sparse 100k step's launch shapes: an empty
 *                                    kernel, a 7.2 MB producer, a consumer that
 *                                    reads it and writes 12.8 MB, and the two
 *                                    as a dependent pair; then an 8 MiB
 *                                    hand-over between two kernels whose
 *                                    workgroup b reads what workgroup b + s
 *                                    wrote, on the same XCD (s = 0, 8) or on
 *                                    another (s = 1, 3, 4)
 *   infoli_calib.x sync              hand-over of partial sums inside one kernel,
 *                                    in the sparse 100k step's shape: 245
 *                                    workgroups of 1024 threads in 49 groups
 *                                    of 5, each writing 2048 (sum,
 *                                    compensation) pairs and 32 KB of other
 *                                    state; the last of a group to arrive
 *                                    merges the group's pairs.  Against the
 *                                    same work as two kernels per step, with
 *                                    (1) one agent-scope release / acquire per
 *                                    workgroup, (2) device-coherent atomic
 *                                    loads and stores of the pairs and no
 *                                    fence, (3) a release per wave; counts
 *                                    merged pairs that were not the current
 *                                    step's.  Then a barrier of all the
 *                                    workgroups (one per CU), and of groups of
 *                                    16, 8, 5 and 2 workgroups that wait only
 *                                    for each other, with the same three
 *                                    publication schemes.
 *
 * Every index is drawn in range on the host; grids are sized to the buffers.
 */
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#define CHK(x)                                                                    \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));               \
            exit(1);                                                              \
        }                                                                         \
    } while (0)

__global__ void copy_k(const float4 *__restrict__ a, float4 *__restrict__ b, long n)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long s = (long)gridDim.x * blockDim.x;
    for (; i < n; i += s) b[i] = a[i];
}

/* read-only stream: 4 independent 16-byte loads in flight per thread */
__global__ void read_k(const float4 *__restrict__ a, float *out, long n)
{
    const long s = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0.f;
    for (; i + 3 * s < n; i += 4 * s) {
        const float4 x0 = a[i], x1 = a[i + s], x2 = a[i + 2 * s], x3 = a[i + 3 * s];
        acc += (x0.x + x1.y) + (x2.z + x3.w);
    }
    for (; i < n; i += s) acc += a[i].x;
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

/* copy with 4 independent 16-byte loads in flight per thread */
__global__ void copy4_k(const float4 *__restrict__ a, float4 *__restrict__ b, long n)
{
    const long s = (long)gridDim.x * blockDim.x;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * s < n; i += 4 * s) {
        const float4 x0 = a[i], x1 = a[i + s], x2 = a[i + 2 * s], x3 = a[i + 3 * s];
        b[i] = x0;
        b[i + s] = x1;
        b[i + 2 * s] = x2;
        b[i + 3 * s] = x3;
    }
    for (; i < n; i += s) b[i] = a[i];
}

__global__ void gather_global(const int *__restrict__ idx, const float *__restrict__ v, float *out,
                              int per, long nthreads)
{
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0;
    for (int k = 0; k < per; k++) acc += v[idx[tid + (long)k * nthreads]];
    out[tid] = acc;
}

__global__ void gather_lds(const int *__restrict__ idx, const float *__restrict__ v, int nv, float *out,
                           int per, long nthreads)
{
    extern __shared__ float sv[];
    for (int i = threadIdx.x; i < nv; i += blockDim.x) sv[i] = v[i];
    __syncthreads();
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    float acc = 0;
    for (int k = 0; k < per; k++) acc += sv[idx[tid + (long)k * nthreads]];
    out[tid] = acc;
}

__global__ void exp_rate(float *out, int iters)
{
    float a = threadIdx.x * 1e-3f, b = a + 0.5f, c = a + 0.25f, d = a + 0.125f;
    for (int i = 0; i < iters; i++) {
        a = __builtin_amdgcn_exp2f(a) * -0.5f;
        b = __builtin_amdgcn_exp2f(b) * -0.5f;
        c = __builtin_amdgcn_exp2f(c) * -0.5f;
        d = __builtin_amdgcn_exp2f(d) * -0.5f;
    }
    out[(long)blockIdx.x * blockDim.x + threadIdx.x] = a + b + c + d;
}

/* Launch-cost kernels.  The producer has the tile kernel's shape (441
 * workgroups of 1024 threads, one 16-byte partial per thread); the consumer
 * has the compartment kernel's (256-thread groups over 100k cells' 128 bytes
 * of state). */
__global__ __launch_bounds__(1024) void launch_empty(float4 *p)
{
    if (p == nullptr && threadIdx.x == 12345) p[0] = float4{};
}

__global__ __launch_bounds__(1024) void launch_produce(float4 *p)
{
    const size_t i = (size_t)blockIdx.x * 1024 + threadIdx.x;
    p[i] = float4{1.f, 2.f, 3.f, (float)i};
}

__global__ __launch_bounds__(256) void launch_consume(const float4 *p, float4 *st, size_t np, size_t ns)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const float4 a = i < np ? p[i] : float4{};
    if (i < ns) st[i] = a;
}

/* hand-over: workgroup b writes, then reads, chunk b of 1024 floats */
__global__ __launch_bounds__(256) void handover_write(float *p)
{
    const size_t base = (size_t)blockIdx.x * 1024;
    for (int i = threadIdx.x; i < 1024; i += 256) p[base + i] = (float)(base + i);
}

__global__ __launch_bounds__(256) void handover_read(const float *p, float *q, int shift, int nb)
{
    const size_t src = (size_t)((blockIdx.x + shift) % nb) * 1024, dst = (size_t)blockIdx.x * 1024;
    for (int i = threadIdx.x; i < 1024; i += 256) q[dst + i] = p[src + i] + 1.0f;
}

/* ---- in-kernel hand-over (sync mode) ---------------------------------- */

constexpr int kSyncR = 2048;        // pairs per workgroup (receivers of a tile)
constexpr int kSyncState = 8192;    // other floats each workgroup writes

__global__ void sync_bump(int *it) { *it += 1; }

/* one (sum, compensation) pair as a 64-bit word */
__device__ __forceinline__ unsigned long long pack2(float a, float b)
{
    return (unsigned long long)__float_as_uint(a) | ((unsigned long long)__float_as_uint(b) << 32);
}

template <int V>
__device__ __forceinline__ void sync_store(float2 *p, size_t i, float2 v)
{
    if constexpr (V == 2)
        __hip_atomic_store(reinterpret_cast<unsigned long long *>(p) + i, pack2(v.x, v.y), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    else
        p[i] = v;
}

template <int V>
__device__ __forceinline__ float2 sync_load(const float2 *p, size_t i)
{
    if constexpr (V == 2) {
        const unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p) + i,
                                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_float2(__uint_as_float((unsigned)u), __uint_as_float((unsigned)(u >> 32)));
    }
    return p[i];
}

/* the consumer's work for group g: merge the group's pairs, check them */
template <int V>
__device__ __forceinline__ void sync_merge(const float2 *part, float *out, int g, int group, int it, unsigned *err)
{
    unsigned bad = 0;
    for (int r = threadIdx.x; r < kSyncR; r += blockDim.x) {
        float acc = 0.f;
        for (int k = 0; k < group; k++) {
            const int wg = g * group + k;
            const float2 v = sync_load<V>(part, (size_t)wg * kSyncR + r);
            bad += (v.x != (float)(wg * kSyncR + r) || v.y != (float)it) ? 1u : 0u;
            acc += v.x + v.y;
        }
        out[(size_t)g * kSyncR + r] = acc;
    }
    if (bad) atomicAdd(err, bad);
}

/* V = 0: producer only (the consumer is the next kernel); 1: agent-scope
 * release / acquire once per workgroup; 2: device-coherent atomics, no
 * fence; 3: an agent-scope release in every wave */
template <int V>
__global__ __launch_bounds__(1024) void sync_produce(float2 *part, float *state, float *out, unsigned *cnt,
                                                     const int *it_p, int group, unsigned *err)
{
    const int wg = blockIdx.x, g = wg / group, it = *it_p;
    for (int r = threadIdx.x; r < kSyncR; r += 1024)
        sync_store<V>(part, (size_t)wg * kSyncR + r, make_float2((float)(wg * kSyncR + r), (float)it));
    for (int i = threadIdx.x; i < kSyncState; i += 1024) state[(size_t)wg * kSyncState + i] += 1.0f;
    if constexpr (V == 0) return;
    __shared__ int s_last;
    if constexpr (V == 3) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned old;
        if constexpr (V == 1)
            old = __hip_atomic_fetch_add(&cnt[g], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        else
            old = __hip_atomic_fetch_add(&cnt[g], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const int last = old == (unsigned)(group - 1);
        if (last && V != 2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (last) __hip_atomic_store(&cnt[g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        s_last = last;
    }
    __syncthreads();
    if (s_last) sync_merge<V>(part, out, g, group, it, err);
}

__global__ __launch_bounds__(1024) void sync_consume(const float2 *part, float *out, const int *it_p, int group,
                                                     unsigned *err)
{
    sync_merge<0>(part, out, blockIdx.x, group, *it_p, err);
}

/* grid barrier over gridDim.x co-resident workgroups: every workgroup
 * publishes a 4 KB slice, waits for all, then reads its neighbour's; V as
 * above.  The wait is bounded (a timeout is counted and the loop goes on) */
template <int V>
__global__ __launch_bounds__(1024) void sync_barrier(float *slab, unsigned *cnt_all, int iters, unsigned *err, int group)
{
    __shared__ int s_ok;
    // workgroups [g * group, (g + 1) * group) synchronise among themselves on
    // cnt_all[g * 16] (their own 64-byte line); group = gridDim.x is the grid barrier
    const int wg = blockIdx.x, nb = group, g = blockIdx.x / group, wl = blockIdx.x % group;
    unsigned *cnt = cnt_all + (size_t)g * 16;
    slab += (size_t)g * 2 * group * 1024;
    unsigned bad = 0;
    for (int k = 0; k < iters; k++) {
        float *mine = slab + (size_t)(k & 1) * nb * 1024 + (size_t)wl * 1024;
        const float val = (float)(k * nb + wl + wg * 0);
        if constexpr (V == 2) {
            __hip_atomic_store(mine + threadIdx.x, val, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            mine[threadIdx.x] = val;
        }
        if constexpr (V == 3) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned target = (unsigned)(k + 1) * nb;
            if constexpr (V == 1)
                __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            else
                __hip_atomic_fetch_add(cnt, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            unsigned spins = 0;
            int ok = 1;
            while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++spins > (1u << 22)) { ok = 0; break; }
                __builtin_amdgcn_s_sleep(1);
            }
            if (V != 2) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            s_ok = ok;
        }
        __syncthreads();
        if (!s_ok && threadIdx.x == 0) bad += 1u << 16;
        const float *theirs = slab + (size_t)(k & 1) * nb * 1024 + (size_t)((wl + 1) % nb) * 1024;
        float v;
        if constexpr (V == 2) v = __hip_atomic_load(theirs + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else v = theirs[threadIdx.x];
        bad += v != (float)(k * nb + (wl + 1) % nb) ? 1u : 0u;
    }
    if (bad) atomicAdd(err, bad);
}

static float time_ms(hipEvent_t e0, hipEvent_t e1)
{
    float ms = 0;
    CHK(hipEventSynchronize(e1));
    CHK(hipEventElapsedTime(&ms, e0, e1));
    return ms;
}

int main(int argc, char **argv)
{
    const char *what = argc > 1 ? argv[1] : "copy";
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    if (!strcmp(what, "copy")) {
        const long mib = argc > 2 ? atol(argv[2]) : 1024;
        const long n4 = mib * (1L << 20) / 16;
        float4 *a, *b;
        CHK(hipMalloc(&a, n4 * 16));
        CHK(hipMalloc(&b, n4 * 16));
        CHK(hipMemset(a, 0, n4 * 16));
        for (int rep = 0; rep < 3; rep++) {
            CHK(hipEventRecord(e0));
            hipLaunchKernelGGL(copy_k, dim3(256 * 16), dim3(256), 0, 0, a, b, n4);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            const float ms = time_ms(e0, e1);
            printf("copy %ld MiB: %.3f ms, %.1f GB/s read+write\n", mib, ms, 2.0 * n4 * 16 / ms / 1e6);
        }
        float *red;
        CHK(hipMalloc(&red, 256L * 16 * 256 * sizeof(float)));
        for (int rep = 0; rep < 3; rep++) {
            CHK(hipEventRecord(e0));
            hipLaunchKernelGGL(copy4_k, dim3(256 * 16), dim3(256), 0, 0, a, b, n4);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            const float ms = time_ms(e0, e1);
            printf("copy4 %ld MiB: %.3f ms, %.1f GB/s read+write\n", mib, ms, 2.0 * n4 * 16 / ms / 1e6);
        }
        for (int rep = 0; rep < 3; rep++) {
            CHK(hipEventRecord(e0));
            hipLaunchKernelGGL(read_k, dim3(256 * 16), dim3(256), 0, 0, a, red, n4);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            const float ms = time_ms(e0, e1);
            printf("read %ld MiB: %.3f ms, %.1f GB/s read\n", mib, ms, 1.0 * n4 * 16 / ms / 1e6);
        }
        CHK(hipFree(red));
        CHK(hipFree(a));
        CHK(hipFree(b));
    } else if (!strcmp(what, "gather")) {
        const long nv = argc > 2 ? atol(argv[2]) : 100000;
        const long nthreads = 1L << 20;
        const int per = 16;
        const long ne = nthreads * per;
        std::vector<int> hidx(ne);
        std::mt19937 rng(1);
        // "xcd": workgroup b reads only slice b % 8 of the table.  Workgroups
        // are dealt to the 8 XCDs round-robin, so each XCD's L2 sees one
        // eighth of the table.
        const bool xcd = argc > 3 && !strcmp(argv[3], "xcd") && nv >= 8;
        const long slice = xcd ? nv / 8 : nv;
        std::uniform_int_distribution<int> U(0, (int)slice - 1);
        for (long i = 0; i < ne; i++)
            hidx[i] = U(rng) + (xcd ? (int)(((i % nthreads) / 256) % 8 * slice) : 0);
        std::vector<float> hv(nv, 1.0f);
        int *idx;
        float *v, *out;
        CHK(hipMalloc(&idx, ne * 4));
        CHK(hipMalloc(&v, nv * 4));
        CHK(hipMalloc(&out, nthreads * 4));
        CHK(hipMemcpy(idx, hidx.data(), ne * 4, hipMemcpyHostToDevice));
        CHK(hipMemcpy(v, hv.data(), nv * 4, hipMemcpyHostToDevice));
        for (int rep = 0; rep < 2; rep++) {
            CHK(hipEventRecord(e0));
            hipLaunchKernelGGL(gather_global, dim3(nthreads / 256), dim3(256), 0, 0, idx, v, out, per, nthreads);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            const float ms = time_ms(e0, e1);
            if (rep)
                printf("gather global, table %ld floats%s: %.3f ms, %.1f G gathers/s\n", nv,
                       xcd ? " (one slice per XCD)" : "", ms, ne / ms / 1e6);
        }
        if (nv * 4 <= 64 * 1024) {
            for (int rep = 0; rep < 2; rep++) {
                CHK(hipEventRecord(e0));
                hipLaunchKernelGGL(gather_lds, dim3(nthreads / 256), dim3(256), nv * 4, 0, idx, v, (int)nv, out,
                                   per, nthreads);
                CHK(hipGetLastError());
                CHK(hipEventRecord(e1));
                const float ms = time_ms(e0, e1);
                if (rep) printf("gather LDS, table %ld floats: %.3f ms, %.1f G gathers/s\n", nv, ms, ne / ms / 1e6);
            }
        }
    } else if (!strcmp(what, "exp")) {
        const long nthreads = 1L << 20;
        const int iters = 4096;
        float *out;
        CHK(hipMalloc(&out, nthreads * 4));
        for (int rep = 0; rep < 2; rep++) {
            CHK(hipEventRecord(e0));
            hipLaunchKernelGGL(exp_rate, dim3(nthreads / 256), dim3(256), 0, 0, out, iters);
            CHK(hipGetLastError());
            CHK(hipEventRecord(e1));
            const float ms = time_ms(e0, e1);
            if (rep) printf("exp2 (+1 mul each): %.1f G exp/s\n", (double)nthreads * iters * 4 / ms / 1e6);
        }
    } else if (!strcmp(what, "launch")) {
        const int G = 441, K = 200;
        const size_t np = (size_t)G * 1024, ns = 100000 * 8;
        const dim3 cg((unsigned)((ns + 255) / 256));
        float4 *part, *st;
        CHK(hipMalloc(&part, np * 16));
        CHK(hipMalloc(&st, ns * 16));
        CHK(hipMemset(part, 0, np * 16));
        hipStream_t s;
        CHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        const char *names[] = {"empty 441x1024", "producer 7.2 MB", "consumer 7.2 MB in, 12.8 MB out",
                               "producer -> consumer pair"};
        for (int v = 0; v < 4; v++) {
            hipGraph_t g;
            hipGraphExec_t ge;
            CHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; k++) {
                if (v == 0) hipLaunchKernelGGL(launch_empty, dim3(G), dim3(1024), 0, s, part);
                if (v == 1 || v == 3) hipLaunchKernelGGL(launch_produce, dim3(G), dim3(1024), 0, s, part);
                if (v >= 2) hipLaunchKernelGGL(launch_consume, cg, dim3(256), 0, s, (const float4 *)part, st, np, ns);
            }
            CHK(hipStreamEndCapture(s, &g));
            CHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            float best = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                CHK(hipGraphLaunch(ge, s));
                CHK(hipStreamSynchronize(s));
                CHK(hipEventRecord(e0, s));
                CHK(hipGraphLaunch(ge, s));
                CHK(hipEventRecord(e1, s));
                const float ms = time_ms(e0, e1);
                best = ms < best ? ms : best;
            }
            printf("graph of %d x %-34s %.2f us each\n", K, names[v], best * 1000 / K);
            CHK(hipGraphExecDestroy(ge));
            CHK(hipGraphDestroy(g));
        }
        // hand-over of 8 MiB: 2048 workgroups, each reading the chunk that
        // workgroup b + shift wrote (workgroups are dealt to the 8 XCDs
        // round-robin, so shifts that are multiples of 8 stay on one XCD)
        const int nb = 2048;
        float *hp, *hq;
        CHK(hipMalloc(&hp, (size_t)nb * 4096));
        CHK(hipMalloc(&hq, (size_t)nb * 4096));
        CHK(hipMemset(hp, 0, (size_t)nb * 4096));
        for (int shift : {0, 8, 1, 3, 4}) {
            hipGraph_t g;
            hipGraphExec_t ge;
            CHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; k++) {
                hipLaunchKernelGGL(handover_write, dim3(nb), dim3(256), 0, s, hp);
                hipLaunchKernelGGL(handover_read, dim3(nb), dim3(256), 0, s, (const float *)hp, hq, shift, nb);
            }
            CHK(hipStreamEndCapture(s, &g));
            CHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            float best = 1e30f;
            for (int rep = 0; rep < 3; rep++) {
                CHK(hipGraphLaunch(ge, s));
                CHK(hipStreamSynchronize(s));
                CHK(hipEventRecord(e0, s));
                CHK(hipGraphLaunch(ge, s));
                CHK(hipEventRecord(e1, s));
                const float ms = time_ms(e0, e1);
                best = ms < best ? ms : best;
            }
            printf("8 MiB hand-over, reader b reads writer b + %d (%s XCD): %.2f us per pair\n", shift,
                   shift % 8 ? "another" : "the same", best * 1000 / K);
            CHK(hipGraphExecDestroy(ge));
            CHK(hipGraphDestroy(g));
        }
        CHK(hipFree(hp));
        CHK(hipFree(hq));
        CHK(hipStreamDestroy(s));
        CHK(hipFree(part));
        CHK(hipFree(st));
    } else if (!strcmp(what, "sync")) {
        const int group = 5, G = 49 * group, K = 100;
        float2 *part;
        float *state, *out, *slab;
        unsigned *cnt, *err;
        int *it;
        CHK(hipMalloc(&part, (size_t)G * kSyncR * sizeof(float2)));
        CHK(hipMalloc(&state, (size_t)G * kSyncState * sizeof(float)));
        CHK(hipMalloc(&out, (size_t)(G / group) * kSyncR * sizeof(float)));
        CHK(hipMalloc(&cnt, 64 * sizeof(unsigned)));
        CHK(hipMalloc(&err, sizeof(unsigned)));
        CHK(hipMalloc(&it, sizeof(int)));
        hipStream_t s;
        CHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        CHK(hipMemsetAsync(part, 0, (size_t)G * kSyncR * sizeof(float2), s));
        CHK(hipMemsetAsync(state, 0, (size_t)G * kSyncState * sizeof(float), s));
        CHK(hipMemsetAsync(cnt, 0, 64 * sizeof(unsigned), s));
        CHK(hipMemsetAsync(it, 0, sizeof(int), s));
        const char *names[] = {"two kernels (produce, then merge)", "one kernel, release/acquire per workgroup",
                               "one kernel, device-coherent atomics, no fence", "one kernel, release per wave"};
        for (int v = 0; v < 4; v++) {
            CHK(hipMemsetAsync(err, 0, sizeof(unsigned), s));
            CHK(hipStreamSynchronize(s));
            hipGraph_t g;
            hipGraphExec_t ge;
            CHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < K; k++) {
                hipLaunchKernelGGL(sync_bump, dim3(1), dim3(1), 0, s, it);
                if (v == 0) {
                    hipLaunchKernelGGL(sync_produce<0>, dim3(G), dim3(1024), 0, s, part, state, out, cnt,
                                       (const int *)it, group, err);
                    hipLaunchKernelGGL(sync_consume, dim3(G / group), dim3(1024), 0, s, (const float2 *)part, out,
                                       (const int *)it, group, err);
                } else if (v == 1) {
                    hipLaunchKernelGGL(sync_produce<1>, dim3(G), dim3(1024), 0, s, part, state, out, cnt,
                                       (const int *)it, group, err);
                } else if (v == 2) {
                    hipLaunchKernelGGL(sync_produce<2>, dim3(G), dim3(1024), 0, s, part, state, out, cnt,
                                       (const int *)it, group, err);
                } else {
                    hipLaunchKernelGGL(sync_produce<3>, dim3(G), dim3(1024), 0, s, part, state, out, cnt,
                                       (const int *)it, group, err);
                }
            }
            CHK(hipStreamEndCapture(s, &g));
            CHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            float best = 1e30f;
            for (int rep = 0; rep < 4; rep++) {
                CHK(hipEventRecord(e0, s));
                CHK(hipGraphLaunch(ge, s));
                CHK(hipEventRecord(e1, s));
                const float ms = time_ms(e0, e1);
                best = ms < best ? ms : best;
            }
            unsigned nerr = 0;
            CHK(hipMemcpy(&nerr, err, sizeof nerr, hipMemcpyDeviceToHost));
            printf("hand-over, %-46s %.2f us per step, %u stale pairs of %lld merged\n", names[v],
                   best * 1000 / K, nerr, (long long)4 * K * G * kSyncR);
            CHK(hipGraphExecDestroy(ge));
            CHK(hipGraphDestroy(g));
        }
        // grid barrier, one 1024-thread workgroup per CU (all co-resident)
        int ncu = 0;
        CHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0));
        const int iters = 200;
        CHK(hipMalloc(&slab, (size_t)2 * ncu * 1024 * sizeof(float)));
        unsigned *gcnt;
        CHK(hipMalloc(&gcnt, (size_t)ncu * 16 * sizeof(unsigned)));
        const char *bnames[] = {"", "release/acquire per workgroup", "device-coherent atomics, no fence",
                                "release per wave"};
        // the grid barrier, then groups of workgroups that synchronise only
        // among themselves (all groups at once, ncu workgroups in all)
        for (int group : {ncu, 16, 8, 5, 2}) {
            const int grid = ncu / group * group;
            for (int v = 1; v < 4; v++) {
                CHK(hipMemsetAsync(gcnt, 0, (size_t)ncu * 16 * sizeof(unsigned), s));
                CHK(hipMemsetAsync(err, 0, sizeof(unsigned), s));
                CHK(hipStreamSynchronize(s));
                CHK(hipEventRecord(e0, s));
                if (v == 1) hipLaunchKernelGGL(sync_barrier<1>, dim3(grid), dim3(1024), 0, s, slab, gcnt, iters, err, group);
                if (v == 2) hipLaunchKernelGGL(sync_barrier<2>, dim3(grid), dim3(1024), 0, s, slab, gcnt, iters, err, group);
                if (v == 3) hipLaunchKernelGGL(sync_barrier<3>, dim3(grid), dim3(1024), 0, s, slab, gcnt, iters, err, group);
                CHK(hipEventRecord(e1, s));
                const float ms = time_ms(e0, e1);
                unsigned nerr = 0;
                CHK(hipMemcpy(&nerr, err, sizeof nerr, hipMemcpyDeviceToHost));
                printf("barrier of %3d workgroups (%3d groups), %-34s %.2f us per barrier, %u stale reads, %u timeouts\n",
                       group, grid / group, bnames[v], ms * 1000 / iters, nerr & 0xFFFF, nerr >> 16);
            }
        }
        CHK(hipFree(gcnt));
        CHK(hipFree(slab));
        CHK(hipStreamDestroy(s));
        CHK(hipFree(part));
        CHK(hipFree(state));
        CHK(hipFree(out));
        CHK(hipFree(cnt));
        CHK(hipFree(err));
        CHK(hipFree(it));
    } else {
        fprintf(stderr, "usage: infoli_calib.x copy [MiB] | gather [table] [xcd] | exp | launch | sync\n");
        return 2;
    }
    return 0;
}
