// Device kernels of the gfx950 sync-free lower-triangular solve.
//
// Execution model shared by every solve kernel:
//  * Work is handed out by an atomic ticket, one atomicAdd per workgroup
//    (CUDA/ourWrtFst/spts_syncfree_cuda_csr_wrt.h:24 takes one per thread; a
//    single counter serialises at ~11 ns per increment on MI355X, so per-thread
//    or per-wavefront tickets alone would cost 0.1-0.4 ms on 2M rows).  The
//    grid is persistent: a workgroup takes tickets until the rows run out.
//    Tickets are only taken by running workgroups, so every row a workgroup
//    waits for belongs to a workgroup that is running or finished; that is
//    what makes spinning safe without a grid barrier.
//  * Readiness of x[c] is observed either through the value itself (SYNC
//    VALUE: the readiness array -- x, or a staged copy in uncached memory --
//    starts as a signalling-NaN sentinel that arithmetic can never produce; a
//    relaxed agent-scope 32/64-bit atomic load both tests and fetches it) or
//    through an int flag per row holding the solve's epoch (SYNC FLAG: the
//    value is published with a returning agent-scope exchange and the flag
//    store depends on it; readers load the value with an agent-scope atomic
//    after seeing the flag -- no cache-wide release/acquire fences; the
//    epoch makes resets between solves unnecessary).
//  * Thread-per-row kernels use the write-first loop of
//    spts_syncfree_cuda_csr_wrt.h:37-55: a lane publishes x[i] from inside the
//    loop in which it polls, so lanes of one wavefront that depend on each
//    other always make progress.  Each lane polls a small window of its next
//    dependencies at once (two in the value protocol, kWin with flags) and
//    consumes the ready prefix in stored order, so the accumulation order is
//    the sequential one (bit-exact against the oracle's fused forward
//    substitution).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace cap {

constexpr int kWave = 64;
constexpr int kBlock = 256;       // 4 wavefronts per workgroup
constexpr int kBlockSmall = 128;  // the thread-per-row kernels' other workgroup size
// Thread-per-row workgroups of several right-hand sides (value protocol, no
// LDS window): a ticket covers BLOCK / lanes-per-row rows, and each ticket
// is one atomic on the one ticket counter, which serializes at about 12 ns
// an atomic; 16 right-hand sides give 16 rows a ticket at 256 threads.
constexpr int kBlockLarge = 1024;
constexpr int kWin = 4;      // dependencies polled concurrently per lane (flag protocol)


template <typename T>
struct VBits;
template <>
struct VBits<double> {
    typedef unsigned long long U;
    static constexpr U kSentinel = 0xFFF0000000000001ULL;  // negative signalling NaN
    static constexpr U kQNaN = 0x7FF8000000000000ULL;
};
template <>
struct VBits<float> {
    typedef unsigned int U;
    static constexpr U kSentinel = 0xFF800001u;
    static constexpr U kQNaN = 0x7FC00000u;
};

template <typename U>
__device__ __forceinline__ U ld_agent(const U *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename U>
__device__ __forceinline__ void st_agent(U *p, U v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double as_val(unsigned long long u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ float as_val(unsigned int u) { return __uint_as_float(u); }
__device__ __forceinline__ unsigned long long as_bits(double v) { return (unsigned long long)__double_as_longlong(v); }
__device__ __forceinline__ unsigned int as_bits(float v) { return __float_as_uint(v); }
__device__ __forceinline__ double fmac(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmac(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
// hardware FP / int atomic add at agent scope (global_atomic_add_f64 / _f32)
template <typename V>
__device__ __forceinline__ void atomic_add_agent(V *p, V v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Readiness protocol state shared by the solve kernels.
//  flags/epoch: SYNC FLAG only.  spin_limit / wait_ticks: polls without
//  progress / wall time without progress after which a lane gives up,
//  publishes NaN (so every dependent row, and with it the whole grid, still
//  finishes) and sets bit 0 of *err; <= 0 = no bound of that kind.
//  ready: SYNC VALUE only -- the array that is polled and published (same
//  shape as x, sentinel-filled before the solve); results are also stored
//  into x with plain stores.  nullptr = x itself is polled.  The library
//  places it in uncached device memory, where a poll goes straight to the
//  memory side instead of through an L2 that agent-scope loads must miss.
//  unfused: accumulate s + v*x as a rounded product and a rounded sum (the
//  arithmetic of the reference's serial executor,
//  openCL/SynFree_csc/sptrsv_syncfree_serialref.h:67-68, compiled without
//  contraction) instead of one fused multiply-add.
struct SyncCtx {
    int *flags;
    int epoch;
    int spin_limit;
    int *err;
    void *ready;
    int unfused;
    int wait_ticks;  // wall-time bound of one wait, in 10 ns ticks (<= 0: none)
    // kPollCounters counters of the polls the value-protocol write-first
    // solve issued (its counting instantiation only; nullptr otherwise)
    unsigned long long *polls;
};
constexpr int kPollCounters = 64;

// One accumulation step of a row's sum, in the plan's arithmetic.  The
// operand order and the stored-order chain are the same in both forms; only
// the rounding of the product differs.
template <typename T>
__device__ __forceinline__ T acc(const SyncCtx &sc, T v, T x, T s)
{
    return sc.unfused ? s + v * x : fmac(v, x, s);
}

template <typename T>
__device__ __forceinline__ typename VBits<T>::U *ready_array(const SyncCtx &sc, T *x)
{
    typedef typename VBits<T>::U U;
    return sc.ready ? reinterpret_cast<U *>(sc.ready) : reinterpret_cast<U *>(x);
}
template <typename T>
__device__ __forceinline__ const typename VBits<T>::U *ready_array(const SyncCtx &sc, const T *x)
{
    typedef typename VBits<T>::U U;
    return sc.ready ? reinterpret_cast<const U *>(sc.ready) : reinterpret_cast<const U *>(x);
}

// Watchdog state of one wait: polls since the last progress and the
// real-time counter (100 MHz, s_memrealtime) read at the kWatchCheck-th of
// them.  A wait gives up after sc.spin_limit polls (when > 0) or once
// sc.wait_ticks of wall time have passed since that read (when > 0), so a
// kernel drains in bounded wall time however long each poll takes under
// memory contention.  The counter is read once every kWatchCheck polls.
constexpr unsigned kWatchCheck = 1024;
struct Idle {
    unsigned n = 0, t0 = 0;
    __device__ __forceinline__ void progress() { n = 0; }
};
__device__ __forceinline__ bool expired(const SyncCtx &sc, Idle &w)
{
    ++w.n;
    if (sc.spin_limit > 0 && w.n > (unsigned)sc.spin_limit) return true;
    if ((w.n & (kWatchCheck - 1)) != 0 || sc.wait_ticks <= 0) return false;
    const unsigned now = (unsigned)__builtin_amdgcn_s_memrealtime();
    if (w.n == kWatchCheck) {
        w.t0 = now;
        return false;
    }
    return now - w.t0 > (unsigned)sc.wait_ticks;
}
__device__ __forceinline__ bool give_up(const SyncCtx &sc, Idle &w)
{
    if (!expired(sc, w)) return false;
    atomicOr(sc.err, 1);
    return true;
}

// Persistent ticket loop: the workgroup takes `per_block` work items at a time
// (one atomicAdd, broadcast through LDS), runs init(base) on every thread
// between two barriers (all threads have left the previous body), then
// body(base), until `total` items are handed out.  body must not contain
// workgroup barriers.
template <typename FI, typename F>
__device__ __forceinline__ void ticket_loop(unsigned *ticket, int per_block, int total, FI init, F body)
{
    __shared__ int s_base;
    while (true) {
        if (threadIdx.x == 0) s_base = (int)atomicAdd(ticket, (unsigned)per_block);
        __syncthreads();
        const int base = s_base;
        init(base);
        __syncthreads();
        if (base >= total) return;
        body(base);
    }
}
template <typename F>
__device__ __forceinline__ void ticket_loop(unsigned *ticket, int per_block, int total, F body)
{
    ticket_loop(ticket, per_block, total, [](int) {}, body);
}

// Workgroup-local readiness window: values of the rows [lo, lo + len) the
// workgroup is solving, kept in LDS next to their global copies.  A
// dependency inside the window is polled in LDS (one CU, no memory traffic);
// older rows are polled in global memory.  len == 0 disables it.
template <typename U>
struct LdsWin {
    U *s;
    int lo, len, k;
    __device__ __forceinline__ bool has(int c) const { return (unsigned)(c - lo) < (unsigned)len; }
    __device__ __forceinline__ U load(int c, int q) const
    {
        return __hip_atomic_load(s + (c - lo) * k + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ void store(int c, int q, U v) const
    {
        __hip_atomic_store(s + (c - lo) * k + q, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
};

// Publish x[i] in the value protocol: the sentinel can never be published
// (division results are never signalling NaNs; the guard keeps that true for
// any input bit pattern).
template <typename T>
__device__ __forceinline__ void publish_value(T *x, long long idx, T xi)
{
    typedef typename VBits<T>::U U;
    U o = as_bits(xi);
    if (o == VBits<T>::kSentinel) o = VBits<T>::kQNaN;
    st_agent(reinterpret_cast<U *>(x) + idx, o);
}

// ------------------------------------------------------------------ fills
template <typename U>
__global__ void __launch_bounds__(kBlock) k_fill(U *__restrict__ p, long long count, U value)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = value;
}

// Per-solve reset: tickets (and the value protocol's sentinel fill of x) in
// one launch, so no memset node is needed when a caller captures solves in a
// graph.
template <typename U>
__global__ void __launch_bounds__(kBlock)
    k_prepare(U *__restrict__ p, long long count, U value, unsigned *__restrict__ ticket, int n_ticket)
{
    if (blockIdx.x == 0 && (int)threadIdx.x < n_ticket) ticket[threadIdx.x] = 0u;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = value;
}

// k_prepare plus zeroed convergence norms (two words): the fused refinement
// sweep's whole reset in one launch.
template <typename U>
__global__ void __launch_bounds__(kBlock)
    k_prepare_norms(U *__restrict__ p, long long count, U value, unsigned *__restrict__ ticket, int n_ticket,
                    unsigned long long *__restrict__ norms)
{
    if (blockIdx.x == 0 && (int)threadIdx.x < n_ticket) ticket[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 2) norms[threadIdx.x] = 0ULL;
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = value;
}

// ------------------------------------------------- write-first row solver
// Publishes x[o] (o = row*k + q) in the protocol SYNC (1 value, 2 flag).
template <typename T, int SYNC>
__device__ __forceinline__ void publish(T *x, const SyncCtx &sc, long long o, T xi)
{
    if (SYNC == 1) {
        if (sc.ready) {
            publish_value(reinterpret_cast<T *>(sc.ready), o, xi);
            x[o] = xi;
        } else {
            publish_value(x, o, xi);
        }
    } else {
        // flag protocol without cache-wide fences: the value goes out as a
        // returning agent-scope exchange, and the flag store depends on the
        // returned word, so the flag is issued only once the value is in place
        typedef typename VBits<T>::U U;
        const U old = __hip_atomic_exchange(reinterpret_cast<U *>(x) + o, as_bits(xi), __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
        int ep = sc.epoch;
        asm volatile("" : "+v"(ep) : "v"(old));
        __hip_atomic_store(sc.flags + o, ep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Blocking wait for x[o] (used where every dependency belongs to an earlier
// wavefront).  Returns the value; NaN if the watchdog fired.
template <typename T, int SYNC>
__device__ __forceinline__ T wait_value(const T *x, const SyncCtx &sc, long long o)
{
    typedef typename VBits<T>::U U;
    const U *xu = reinterpret_cast<const U *>(x);
    Idle idle;
    if (SYNC == 1) {
        const U *xr = ready_array(sc, x);
        U raw;
        while ((raw = ld_agent(xr + o)) == VBits<T>::kSentinel)
            if (give_up(sc, idle)) return as_val(VBits<T>::kQNaN);
        return as_val(raw);
    } else {
        while (ld_agent(sc.flags + o) != sc.epoch)
            if (give_up(sc, idle)) return as_val(VBits<T>::kQNaN);
        return as_val(ld_agent(xu + o));  // issued after the flag was seen
    }
}

// Non-blocking test of x[o]: true and the value when published.
template <typename T, int SYNC>
__device__ __forceinline__ bool try_value(const T *x, const SyncCtx &sc, long long o, T *out)
{
    typedef typename VBits<T>::U U;
    const U *xu = reinterpret_cast<const U *>(x);
    if (SYNC == 1) {
        const U raw = ld_agent(ready_array(sc, x) + o);
        *out = as_val(raw);
        return raw != VBits<T>::kSentinel;
    }
    if (ld_agent(sc.flags + o) != sc.epoch) return false;
    *out = as_val(ld_agent(xu + o));  // issued after the flag was seen
    return true;
}

__device__ __forceinline__ double bcast(double v, int lane)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, lane);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float bcast(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// One row solved by a whole wavefront in the sequential order: lane l polls
// off-diagonal l of the current 64-entry window; the wavefront then adds the
// ready prefix term by term, s = fma(v_t, x_t, s) with broadcast operands, so
// the sum is the oracle's fused sequential one and only the last dependency
// to arrive sits on the critical path (one fma and the division after it).
// Every dependency is an earlier row, solved by an earlier wavefront.
template <typename T, int SYNC>
__device__ __forceinline__ void wave_row(const int *__restrict__ rp, const int *__restrict__ ci,
                                         const T *__restrict__ val, const T *__restrict__ b, T *__restrict__ x,
                                         const SyncCtx &sc, const LdsWin<typename VBits<T>::U> &win, int row)
{
    typedef typename VBits<T>::U U;
    const int lane = threadIdx.x & (kWave - 1);
    const int d = rp[row + 1] - 1;
    const T brow = b[row], vd = val[d];  // loaded now, off the critical path
    T s = (T)0;
    Idle idle;
    bool failed = false;
    for (int j0 = rp[row]; j0 < d; j0 += kWave) {
        const int cnt = min(kWave, d - j0);
        const int c = lane < cnt ? ci[j0 + lane] : 0;
        const T v = lane < cnt ? val[j0 + lane] : (T)0;
        T xv = (T)0;
        bool rdy = lane >= cnt;
        int used = 0;
        while (used < cnt) {
            if (!rdy) {
                if (win.has(c)) {
                    const U raw = win.load(c, 0);
                    rdy = raw != VBits<T>::kSentinel;
                    xv = as_val(raw);
                } else {
                    rdy = try_value<T, SYNC>(x, sc, c, &xv);
                }
            }
            const unsigned long long ready = __ballot(rdy);
            const int prefix = ~ready ? __builtin_ctzll(~ready) : kWave;
            const int upto = min(prefix, cnt);
            for (int t = used; t < upto; ++t) s = acc(sc, bcast(v, t), bcast(xv, t), s);
            if (upto > used) {
                idle.progress();
            } else if (give_up(sc, idle)) {  // uniform: idle is the same in every lane
                failed = true;
                break;
            }
            used = upto;
        }
        if (failed) break;
    }
    if (lane == 0) {
        const T xi = failed ? as_val(VBits<T>::kQNaN) : (brow - s) / vd;
        if (win.len) {
            U o = as_bits(xi);
            if (o == VBits<T>::kSentinel) o = VBits<T>::kQNaN;
            win.store(row, 0, o);
        }
        publish<T, SYNC>(x, sc, row, xi);
    }
}

// Value protocol form of wrtfst_row: a window of the next two dependencies
// (entries j and j+1) is polled at once and slides by the number consumed,
// so most rows (one or two off-diagonals) are covered by a single window and
// the loop body stays short.  Consumption is in stored order (fma chain).
// COUNT: add the number of poll loads issued to *np (the per-row spin
// counter of the counting instantiation).
template <typename T, bool COUNT = false>
__device__ __forceinline__ void wrtfst_row_value(const int *__restrict__ ci, const T *__restrict__ val,
                                                 const T *__restrict__ b, T *__restrict__ x, const SyncCtx &sc,
                                                 int i, int j, int d, int k, int q, T s, int n_rev,
                                                 const LdsWin<typename VBits<T>::U> &win, unsigned *np = nullptr)
{
    typedef typename VBits<T>::U U;
    constexpr U kS = VBits<T>::kSentinel;
    U *xr = ready_array(sc, x);
    // b and the diagonal are loaded before the polling loop: read after the
    // last dependency arrives, they would add a memory round trip per level
    const long long o = (long long)(n_rev ? n_rev - 1 - i : i) * k + q;
    const T bo = b[o], vd = val[d];
    int c0 = 0, c1 = 0;
    T v0 = (T)0, v1 = (T)0;
    U x0 = kS, x1 = kS;
    if (j < d) {
        c0 = n_rev ? n_rev - 1 - ci[j] : ci[j];
        v0 = val[j];
    }
    if (j + 1 < d) {
        c1 = n_rev ? n_rev - 1 - ci[j + 1] : ci[j + 1];
        v1 = val[j + 1];
    }
    Idle idle;
    bool done = false;
    do {  // wave-uniform exit: see wrtfst_row
        if (!done) {
            bool prog = false;
            if (COUNT) *np += (j < d && x0 == kS) + (j + 1 < d && x1 == kS);
            if (j < d && x0 == kS) x0 = win.has(c0) ? win.load(c0, q) : ld_agent(xr + (long long)c0 * k + q);
            if (j + 1 < d && x1 == kS) x1 = win.has(c1) ? win.load(c1, q) : ld_agent(xr + (long long)c1 * k + q);
            if (j < d && x0 != kS) {
                s = acc(sc, v0, as_val(x0), s);
                ++j;
                prog = true;
                if (j < d && x1 != kS) {  // both consumed: a fresh window
                    s = acc(sc, v1, as_val(x1), s);
                    ++j;
                    x0 = kS;
                    if (j < d) {
                        c0 = n_rev ? n_rev - 1 - ci[j] : ci[j];
                        v0 = val[j];
                    }
                } else {  // slide by one
                    c0 = c1;
                    v0 = v1;
                    x0 = x1;
                }
                x1 = kS;
                if (j + 1 < d) {
                    c1 = n_rev ? n_rev - 1 - ci[j + 1] : ci[j + 1];
                    v1 = val[j + 1];
                }
            }
            if (prog) {
                idle.progress();
            } else if (j < d && give_up(sc, idle)) {
                s = as_val(VBits<T>::kQNaN);  // poison this row; it still publishes below
                j = d;
            }
            if (j >= d) {  // write first: publish inside the polling loop
                const T xi = (bo - s) / vd;
                U u = as_bits(xi);
                if (u == kS) u = VBits<T>::kQNaN;
                if (win.len) win.store(i, q, u);
                st_agent(xr + o, u);
                if (sc.ready) x[o] = xi;
                asm volatile("" ::: "memory");
                done = true;
            }
        }
    } while (__any(!done));
}

// Solves row i (off-diagonal entries [j, d), diagonal at d) for right-hand
// side q of k (x and b interleaved: x[row*k + q]) and publishes it.  `s`
// carries a partial sum when the caller already consumed a prefix (two-part).
// n_rev > 0 stores row r at position n_rev-1-r of b and x (the adjoint solves
// L^T in reversed row order, which is lower triangular).
template <typename T, int SYNC, bool COUNT = false>
__device__ __forceinline__ void wrtfst_row(const int *__restrict__ ci, const T *__restrict__ val,
                                           const T *__restrict__ b, T *__restrict__ x, const SyncCtx &sc,
                                           int i, int j, int d, int k, int q, T s, int n_rev = 0,
                                           LdsWin<typename VBits<T>::U> win = {nullptr, 0, 0, 1},
                                           unsigned *np = nullptr)
{
    if constexpr (SYNC == 1) {
        wrtfst_row_value<T, COUNT>(ci, val, b, x, sc, i, j, d, k, q, s, n_rev, win, np);
        return;
    }
    // flag protocol: a fixed window of up to kWin entries, refilled when consumed
    typedef typename VBits<T>::U U;
    U *xu = reinterpret_cast<U *>(x);
    const long long o = (long long)(n_rev ? n_rev - 1 - i : i) * k + q;
    const T bo = b[o], vd = val[d];  // before the loop, off the critical path
    int cw[kWin];
    T vw[kWin];
    U xw[kWin];
    bool rdy[kWin];
    int cnt = min(kWin, d - j), used = 0;
    Idle idle;
#pragma unroll
    for (int t = 0; t < kWin; ++t) {
        rdy[t] = false;
        xw[t] = 0;
        cw[t] = t < cnt ? (n_rev ? n_rev - 1 - ci[j + t] : ci[j + t]) : 0;
        vw[t] = t < cnt ? val[j + t] : (T)0;
    }
    // The loop exits only when every lane of the wavefront has published
    // (wave-uniform exit).  With a per-lane exit the compiler may legally
    // move the publishing store to the loop exit, where a finished lane waits
    // for the rest of its wavefront -- which may be waiting for that store.
    bool done = false;
    do {
        if (!done) {
            if (used < cnt) {
                int fl[kWin];
                bool any = false;
#pragma unroll
                for (int t = 0; t < kWin; ++t)
                    fl[t] = (t >= used && t < cnt && !rdy[t]) ? ld_agent(sc.flags + (long long)cw[t] * k + q) : 0;
#pragma unroll
                for (int t = 0; t < kWin; ++t)
                    if (t >= used && t < cnt && !rdy[t] && fl[t] == sc.epoch) any = true;
                if (any) {  // values read after their flags were seen (agent-scope loads)
#pragma unroll
                    for (int t = 0; t < kWin; ++t)
                        if (t >= used && t < cnt && !rdy[t] && fl[t] == sc.epoch) {
                            xw[t] = ld_agent(xu + (long long)cw[t] * k + q);
                            rdy[t] = true;
                        }
                }
                // consume the ready prefix in stored order
                const int before = used;
#pragma unroll
                for (int t = 0; t < kWin; ++t)
                    if (t == used && t < cnt && rdy[t]) {
                        s = acc(sc, vw[t], as_val(xw[t]), s);
                        ++used;
                    }
                if (used != before) {
                    idle.progress();
                } else if (give_up(sc, idle)) {
                    s = as_val(VBits<T>::kQNaN);  // poison this row; it still publishes below
                    used = cnt;
                    j = d;
                }
                if (used == cnt && j < d) {
                    j += cnt;
                    if (j < d) {  // next window
                        cnt = min(kWin, d - j);
                        used = 0;
#pragma unroll
                        for (int t = 0; t < kWin; ++t) {
                            rdy[t] = false;
                            if (t < cnt) {
                                cw[t] = n_rev ? n_rev - 1 - ci[j + t] : ci[j + t];
                                vw[t] = val[j + t];
                            }
                        }
                    }
                }
            }
            if (used == cnt && j >= d) {  // write first: publish inside the polling loop
                publish<T, SYNC>(x, sc, o, (bo - s) / vd);
                asm volatile("" ::: "memory");
                done = true;
            }
        }
    } while (__any(!done));
}

// ourWrtFst (CUDA/ourWrtFst/spts_syncfree_cuda_csr_wrt.h:10-57): one lane
// per row and right-hand side; a workgroup of BLOCK threads (128 or 256)
// takes tickets of BLOCK >> lpr_log2 rows with 1 << lpr_log2 lanes per row,
// lane q of a row owning right-hand side q (lanes with q >= k idle).  The
// workgroup size sets how many rows poll at once on each compute unit.
// sched (optional, n_sched blocks): ticket t solves rows [sched[t].x,
// sched[t].x + sched[t].y) instead of the t-th run of consecutive rows.  The
// host orders the blocks so that every row comes after the rows it depends
// on (each matrix of a batch keeps its row order), which keeps the
// forward-progress argument of the ticket loop.
// COUNT (value protocol only): every lane counts the poll loads it issues;
// each wavefront adds its total to one of sc.polls' kPollCounters counters
// when the grid drains (one atomic per wavefront).
// TILE: a ticket's rows are consecutive, so their entries are one range
// [rp[lo], rp[lo + len]) of col_idx and val; the workgroup first copies it
// into LDS with coalesced loads (consecutive lanes, consecutive entries),
// and the rows read their entries there.  A range longer than kRowTile
// entries (uniform in the workgroup) is read from global memory as usual.
constexpr int kRowTile = 2048;
template <typename T, int SYNC, int LDS, int BLOCK, bool COUNT = false, bool TILE = false>
__global__ void __launch_bounds__(BLOCK)
    k_wrtfst(const int *__restrict__ rp, const int *__restrict__ ci, const T *__restrict__ val,
             const T *__restrict__ b, T *__restrict__ x, SyncCtx sc, int n, int k, int lpr_log2,
             unsigned *__restrict__ ticket, int n_rev, const int2 *__restrict__ sched, int n_sched)
{
    typedef typename VBits<T>::U U;
    static_assert(!COUNT || SYNC == 1, "the poll counter is for the value protocol");
    __shared__ U s_x[LDS ? BLOCK : 1];
    __shared__ int s_ci[TILE ? kRowTile : 1];
    __shared__ T s_v[TILE ? kRowTile : 1];
    const int q = threadIdx.x & ((1 << lpr_log2) - 1);
    const int r = threadIdx.x >> lpr_log2;
    const int rows = BLOCK >> lpr_log2;
    unsigned np = 0;
    ticket_loop(
        ticket, sched ? 1 : rows, sched ? n_sched : n,
        [&](int) {
            if (LDS) s_x[threadIdx.x] = VBits<T>::kSentinel;
        },
        [&](int base) {
            int lo = base, len = rows;
            if (sched) {
                const int2 blk = sched[base];
                lo = blk.x;
                len = blk.y;
            }
            const int i = lo + r;
            LdsWin<U> win{s_x, lo, (LDS && !n_rev) ? len : 0, k};
            const int *cix = ci;
            const T *vx = val;
            int e0 = 0;  // entry j of a row is cix[j - e0]
            if constexpr (TILE) {
                const int first = rp[lo], last = rp[min(lo + len, n)];
                if (last - first <= kRowTile) {  // uniform in the workgroup
                    for (int e = threadIdx.x; e < last - first; e += BLOCK) {
                        s_ci[e] = ci[first + e];
                        s_v[e] = val[first + e];
                    }
                    __syncthreads();
                    cix = s_ci;
                    vx = s_v;
                    e0 = first;
                }
            }
            if (r < len && i < n && q < k)
                wrtfst_row<T, SYNC, COUNT>(cix, vx, b, x, sc, i, rp[i] - e0, rp[i + 1] - 1 - e0, k, q, (T)0, n_rev,
                                           win, &np);
        });
    if constexpr (COUNT) {
        unsigned long long t = np;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) t += (unsigned long long)__shfl_xor((long long)t, o, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0)
            atomic_add_agent(sc.polls + (blockIdx.x * (BLOCK / kWave) + threadIdx.x / kWave) % kPollCounters, t);
    }
}

// Value protocol row solver for R right-hand sides per lane: the lane owns
// right-hand sides q, q + L, ..., q + (R-1)L of row i (those below k), L
// lanes per row, so each poll instruction of the row's L lanes reads L
// contiguous values.  A dependency is consumed once all of the lane's R
// values of it are published; each right-hand side keeps its own sum in
// stored order, so the results are those of R single-RHS lanes.  The window
// (two dependencies, sliding) and the wave-uniform exit are those of
// wrtfst_row_value.
template <typename T, int R>
__device__ __forceinline__ void wrtfst_row_value_r(const int *__restrict__ ci, const T *__restrict__ val,
                                                   const T *__restrict__ b, T *__restrict__ x, const SyncCtx &sc,
                                                   int i, int j, int d, int k, int q, int L)
{
    typedef typename VBits<T>::U U;
    constexpr U kS = VBits<T>::kSentinel;
    U *xr = ready_array(sc, x);
    int nr = 0;  // right-hand sides of this lane
#pragma unroll
    for (int t = 0; t < R; ++t) nr += q + t * L < k;
    const long long o = (long long)i * k + q;
    T bo[R], s[R];
    const T vd = val[d];
#pragma unroll
    for (int t = 0; t < R; ++t) {
        bo[t] = t < nr ? b[o + t * L] : (T)0;  // before the polling loop, as in wrtfst_row_value
        s[t] = (T)0;
    }
    int c0 = 0, c1 = 0;
    T v0 = (T)0, v1 = (T)0;
    U x0[R], x1[R];
#pragma unroll
    for (int t = 0; t < R; ++t) x0[t] = x1[t] = kS;
    if (j < d) {
        c0 = ci[j];
        v0 = val[j];
    }
    if (j + 1 < d) {
        c1 = ci[j + 1];
        v1 = val[j + 1];
    }
    Idle idle;
    bool done = false;
    do {  // wave-uniform exit: see wrtfst_row
        if (!done) {
            bool prog = false, r0 = j < d, r1 = j + 1 < d;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                if (t < nr && j < d && x0[t] == kS) x0[t] = ld_agent(xr + (long long)c0 * k + q + t * L);
                if (t < nr && j + 1 < d && x1[t] == kS) x1[t] = ld_agent(xr + (long long)c1 * k + q + t * L);
                r0 = r0 && (t >= nr || x0[t] != kS);
                r1 = r1 && (t >= nr || x1[t] != kS);
            }
            if (r0) {
#pragma unroll
                for (int t = 0; t < R; ++t) s[t] = acc(sc, v0, as_val(x0[t]), s[t]);
                ++j;
                prog = true;
                if (r1) {  // both consumed: a fresh window
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        s[t] = acc(sc, v1, as_val(x1[t]), s[t]);
                        x0[t] = kS;
                    }
                    ++j;
                    if (j < d) {
                        c0 = ci[j];
                        v0 = val[j];
                    }
                } else {  // slide by one
                    c0 = c1;
                    v0 = v1;
#pragma unroll
                    for (int t = 0; t < R; ++t) x0[t] = x1[t];
                }
#pragma unroll
                for (int t = 0; t < R; ++t) x1[t] = kS;
                if (j + 1 < d) {
                    c1 = ci[j + 1];
                    v1 = val[j + 1];
                }
            }
            if (prog) {
                idle.progress();
            } else if (j < d && give_up(sc, idle)) {
#pragma unroll
                for (int t = 0; t < R; ++t) s[t] = as_val(VBits<T>::kQNaN);  // poison; still published below
                j = d;
            }
            if (j >= d) {  // write first: publish inside the polling loop
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    if (t < nr) {
                        const T xi = (bo[t] - s[t]) / vd;
                        U u = as_bits(xi);
                        if (u == kS) u = VBits<T>::kQNaN;
                        st_agent(xr + o + t * L, u);
                        if (sc.ready) x[o + t * L] = xi;
                    }
                }
                asm volatile("" ::: "memory");
                done = true;
            }
        }
    } while (__any(!done));
}

// Thread-per-row solve with R right-hand sides per lane (value protocol, no
// LDS window): L = 1 << lpr_log2 lanes per row, BLOCK >> lpr_log2 rows per
// ticket.  With k right-hand sides, R per lane makes a ticket R times as
// many rows as k_wrtfst's, every row still polled by its own lanes at once
// (a ticket's rows are all in flight together).  Tickets, batch schedule
// and forward progress as in k_wrtfst.
template <typename T, int BLOCK, int R>
__global__ void __launch_bounds__(BLOCK)
    k_wrtfst_r(const int *__restrict__ rp, const int *__restrict__ ci, const T *__restrict__ val,
               const T *__restrict__ b, T *__restrict__ x, SyncCtx sc, int n, int k, int lpr_log2,
               unsigned *__restrict__ ticket, const int2 *__restrict__ sched, int n_sched)
{
    const int L = 1 << lpr_log2;
    const int q = threadIdx.x & (L - 1);
    const int r = threadIdx.x >> lpr_log2;
    const int rows = BLOCK >> lpr_log2;
    ticket_loop(ticket, sched ? 1 : rows, sched ? n_sched : n, [&](int base) {
        int lo = base, len = rows;
        if (sched) {
            const int2 blk = sched[base];
            lo = blk.x;
            len = blk.y;
        }
        const int i = lo + r;
        if (r < len && i < n && q < k) wrtfst_row_value_r<T, R>(ci, val, b, x, sc, i, rp[i], rp[i + 1] - 1, k, q, L);
    });
}

// ------------------------------------------------------------ refinement
// Maxima of non-negative doubles are taken on their bit patterns (which
// order like the values): per thread over a grid-stride loop, then across
// the workgroup in LDS, then one atomicMax per workgroup -- an atomic per
// element on one address serialises in the memory side.
__device__ __forceinline__ void block_max_to(unsigned long long v, unsigned long long *__restrict__ out)
{
    __shared__ unsigned long long s_max[kBlock / kWave];
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        const unsigned long long w = (unsigned long long)__shfl_xor((long long)v, o, kWave);
        v = w > v ? w : v;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) s_max[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) v = s_max[w] > v ? s_max[w] : v;
        atomicMax(out, v);
    }
}

// ------------------------------------------------ fused refinement sweep
// The first two steps of mixed-precision refinement in one dependency sweep.
// Row i computes, from the pairs (x0_j, d_j) of its dependencies j:
//   x0_i = ((float)b_i - sum v32_j x0_j) * r32_i           FP32 solve of b
//   r_i  = b_i - (sum v_j x0_j + v_ii x0_i)                FP64 residual
//   d_i  = ((float)r_i - sum v32_j d_j) * r32_i            FP32 correction
//   x_i  = (double)x0_i + (double)d_i
// (v32 = (float)v, r32_i = 1.0f / v32_ii computed before the row's polling
// loop, so no division is left on the critical path; the correction absorbs
// the extra rounding of the product; every sum an fma chain in stored order),
// which is bit for bit what the two steps give when run one after the other:
// FP32 solve, residual, FP32 solve of the residual, accumulation.  Row i publishes
// (x0_i, d_i) as one 64-bit word (x0 in the low half) in the value protocol;
// a dependency's word is read once and feeds all three chains, so the sweep
// has the dependency depth of a single solve.  The polling window is the one
// of wrtfst_row_value (two entries, sliding by the number consumed).
constexpr unsigned long long kPairQNaN = 0x7FC000007FC00000ULL;

__device__ __forceinline__ void refine_pair_take(double v, unsigned long long p, float &s0, double &s, float &s1)
{
    const float xa = __uint_as_float((unsigned)p), da = __uint_as_float((unsigned)(p >> 32));
    const float v32 = (float)v;
    s0 = __builtin_fmaf(v32, xa, s0);
    s = __builtin_fma(v, (double)xa, s);
    s1 = __builtin_fmaf(v32, da, s1);
}

__device__ __forceinline__ void refine_pair_row(const int *__restrict__ ci, const double *__restrict__ val,
                                                const double *__restrict__ b, double *__restrict__ x,
                                                unsigned long long *__restrict__ pr, const SyncCtx &sc, int i,
                                                int j, int d, int k, int q,
                                                const LdsWin<unsigned long long> &win, unsigned long long &dmax,
                                                unsigned long long &xmax)
{
    typedef unsigned long long U;
    constexpr U kS = VBits<double>::kSentinel;
    int c0 = 0, c1 = 0;
    double v0 = 0.0, v1 = 0.0;
    U p0 = kS, p1 = kS;
    float s0 = 0.0f, s1 = 0.0f;
    double s = 0.0;
    const long long o = (long long)i * k + q;
    const double vd = val[d], bo = b[o];  // before the loop, off the critical path
    const float rd32 = 1.0f / (float)vd;
    if (j < d) {
        c0 = ci[j];
        v0 = val[j];
    }
    if (j + 1 < d) {
        c1 = ci[j + 1];
        v1 = val[j + 1];
    }
    Idle idle;
    bool done = false;
    do {  // wave-uniform exit: see wrtfst_row
        if (!done) {
            bool prog = false;
            if (j < d && p0 == kS) p0 = win.has(c0) ? win.load(c0, q) : ld_agent(pr + (long long)c0 * k + q);
            if (j + 1 < d && p1 == kS) p1 = win.has(c1) ? win.load(c1, q) : ld_agent(pr + (long long)c1 * k + q);
            if (j < d && p0 != kS) {
                refine_pair_take(v0, p0, s0, s, s1);
                ++j;
                prog = true;
                if (j < d && p1 != kS) {  // both consumed: a fresh window
                    refine_pair_take(v1, p1, s0, s, s1);
                    ++j;
                    p0 = kS;
                    if (j < d) {
                        c0 = ci[j];
                        v0 = val[j];
                    }
                } else {  // slide by one
                    c0 = c1;
                    v0 = v1;
                    p0 = p1;
                }
                p1 = kS;
                if (j + 1 < d) {
                    c1 = ci[j + 1];
                    v1 = val[j + 1];
                }
            }
            if (prog) {
                idle.progress();
            } else if (j < d && give_up(sc, idle)) {
                s0 = s1 = __builtin_nanf("");  // poison this row; it still publishes below
                s = __builtin_nan("");
                j = d;
            }
            if (j >= d) {  // write first: publish inside the polling loop
                const float x0 = ((float)bo - s0) * rd32;
                s = __builtin_fma(vd, (double)x0, s);
                const float dc = ((float)(bo - s) - s1) * rd32;
                U u = ((U)__float_as_uint(dc) << 32) | (U)__float_as_uint(x0);
                if (u == kS) u = kPairQNaN;
                if (win.len) win.store(i, q, u);
                st_agent(pr + o, u);
                const double xd = (double)x0, dd = (double)dc;
                x[o] = xd + dd;
                const U ad = (U)__double_as_longlong(fabs(dd)), ax = (U)__double_as_longlong(fabs(xd));
                dmax = ad > dmax ? ad : dmax;
                xmax = ax > xmax ? ax : xmax;
                asm volatile("" ::: "memory");
                done = true;
            }
        }
    } while (__any(!done));
}

// Thread-per-row launch of the fused sweep (workgroup size, lane layout,
// tickets, batch schedule and LDS window as in k_wrtfst); pr holds the pairs
// and starts filled with the sentinel.  norms[0] = max|d|, norms[1] =
// max|x0| (bits of non-negative doubles, atomicMax per workgroup).  (Row
// tiles, which speed up the batch's 1024-thread solve, gave this sweep
// nothing: 9.60 against 9.45-9.58 ms on 253 matrices x 16 right-hand sides.)
template <int LDS, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
    k_refine2(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
              const double *__restrict__ b, double *__restrict__ x, unsigned long long *__restrict__ pr, SyncCtx sc,
              int n, int k, int lpr_log2, unsigned *__restrict__ ticket, const int2 *__restrict__ sched, int n_sched,
              unsigned long long *__restrict__ norms)
{
    typedef unsigned long long U;
    __shared__ U s_x[LDS ? BLOCK : 1];
    const int q = threadIdx.x & ((1 << lpr_log2) - 1);
    const int r = threadIdx.x >> lpr_log2;
    const int rows = BLOCK >> lpr_log2;
    U dmax = 0, xmax = 0;
    ticket_loop(
        ticket, sched ? 1 : rows, sched ? n_sched : n,
        [&](int) {
            if (LDS) s_x[threadIdx.x] = VBits<double>::kSentinel;
        },
        [&](int base) {
            int lo = base, len = rows;
            if (sched) {
                const int2 blk = sched[base];
                lo = blk.x;
                len = blk.y;
            }
            const int i = lo + r;
            LdsWin<U> win{s_x, lo, LDS ? len : 0, k};
            if (r < len && i < n && q < k)
                refine_pair_row(ci, val, b, x, pr, sc, i, rp[i], rp[i + 1] - 1, k, q, win, dmax, xmax);
        });
    block_max_to(dmax, norms);
    __syncthreads();  // block_max_to's LDS slots are reused by the second reduction
    block_max_to(xmax, norms + 1);
}

// our2Part (CUDA/our2Part/spts_syncfree_cuda_csr.h:10-124): dependencies on
// rows of earlier wavefronts (col < first row of this wave) are waited for
// one by one; the rest is the write-first loop.  Single right-hand side.
template <typename T, int SYNC>
__global__ void __launch_bounds__(kBlock)
    k_twopart(const int *__restrict__ rp, const int *__restrict__ ci, const T *__restrict__ val,
              const T *__restrict__ b, T *__restrict__ x, SyncCtx sc, int n, unsigned *__restrict__ ticket)
{
    ticket_loop(ticket, kBlock, n, [&](int base) {
        const int wave_begin = base + (threadIdx.x & ~(kWave - 1));
        const int i = base + threadIdx.x;
        if (i >= n) return;
        int j = rp[i];
        const int d = rp[i + 1] - 1;
        T s = (T)0;
        for (; j < d; ++j) {  // rows of earlier wavefronts: blocking waits
            const int c = ci[j];
            if (c >= wave_begin) break;
            s = acc(sc, val[j], wait_value<T, SYNC>(x, sc, c), s);
        }
        wrtfst_row<T, SYNC>(ci, val, b, x, sc, i, j, d, 1, 0, s);
    });
}

// mix (CUDA/mix/spts_syncfree_cuda_csr_wrt.h:15-137): unit u covers rows
// [unit[u], unit[u+1]).  A unit of one row is solved by the whole wavefront
// (wave_row: parallel polls, sequential fused sum); a longer unit is solved
// thread-per-row with the write-first loop.  The 16 units of a ticket share
// an LDS window.  Single right-hand side.
// TILE: the ticket's units are consecutive rows [s_lo, s_lo + s_len), whose
// entries are first copied into LDS with coalesced loads (up to kMixTile of
// them; a longer range, uniform in the workgroup, is read from global memory)
// for the thread-per-row units.  Single-row units read global memory.
constexpr int kMixTile = 4096;
template <typename T, int SYNC, int BS, bool TILE = false>
__global__ void __launch_bounds__(BS)
    k_mix(const int *__restrict__ rp, const int *__restrict__ ci, const T *__restrict__ val,
          const T *__restrict__ b, T *__restrict__ x, SyncCtx sc, const int *__restrict__ unit, int n_units,
          unsigned *__restrict__ ticket)
{
    typedef typename VBits<T>::U U;
    constexpr int kUnits = BS / kWave;     // one unit per wavefront per ticket
    __shared__ U s_x[BS / kWave * kWave];  // rows of the ticket: <= 64 per unit
    __shared__ int s_lo, s_len;
    __shared__ int s_ci[TILE ? kMixTile : 1];
    __shared__ T s_v[TILE ? kMixTile : 1];
    const int lane = threadIdx.x & (kWave - 1);
    ticket_loop(
        ticket, kUnits, n_units,
        [&](int base) {
            if (SYNC == 1) s_x[threadIdx.x] = VBits<T>::kSentinel;
            if (threadIdx.x == 0 && base < n_units) {
                s_lo = unit[base];
                s_len = unit[min(base + kUnits, n_units)] - s_lo;
            }
        },
        [&](int base) {
            const int *cix = ci;
            const T *vx = val;
            int e0 = 0;  // entry j of a thread-per-row unit's row is cix[j - e0]
            if constexpr (TILE) {
                const int first = rp[s_lo], last = rp[s_lo + s_len];
                if (last - first <= kMixTile) {  // uniform in the workgroup
                    for (int e = threadIdx.x; e < last - first; e += BS) {
                        s_ci[e] = ci[first + e];
                        s_v[e] = val[first + e];
                    }
                    __syncthreads();
                    cix = s_ci;
                    vx = s_v;
                    e0 = first;
                }
            }
            const int u = base + (int)(threadIdx.x / kWave);
            if (u >= n_units) return;
            const LdsWin<U> win{s_x, s_lo, SYNC == 1 ? s_len : 0, 1};
            const int r0 = unit[u], r1 = unit[u + 1];
            if (r1 - r0 > 1) {
                const int i = r0 + lane;
                if (i < r1)
                    wrtfst_row<T, SYNC>(cix, vx, b, x, sc, i, rp[i] - e0, rp[i + 1] - 1 - e0, 1, 0, (T)0, 0, win);
                return;
            }
            wave_row<T, SYNC>(rp, ci, val, b, x, sc, win, r0);
        });
}

// ---------------------------------------- local: one matrix per workgroup
// For deep matrices (few rows per level) the cross-CU publish -> observe
// round trip (~0.5-2 us) is paid once per level.  Here one workgroup of
// kLocalBlock threads solves a whole matrix -- a segment [lo, hi) of rows
// whose dependencies never leave it -- so every hop stays inside one
// compute unit.  Wavefront w solves the rows lo + w, lo + w + W, ... (W
// wavefronts) one at a time: a row's entries are loaded 64 at a time into
// the lanes (coalesced; the next row's first 64 are prefetched while the
// current row is solved) and broadcast one by one, lane q < k accumulating
// right-hand side q in stored order (fused, bit-exact against the oracle).
//
// Readiness, per segment:
//  * ring (bandwidth + 2 W <= R slots): results are published into an LDS
//    ring -- slot r mod R holds the k values of row r, then a tag r written
//    after them -- and polled there (tag read before the value).  A
//    wavefront publishes row p only once every row <= p - R + bandwidth is
//    done, so no reader of the slot's previous row p - R is left; the
//    frontier is the lowest row any wavefront is still solving.  The
//    lowest unfinished row can always publish (R > bandwidth), so the
//    workgroup always makes progress.
//  * global: x itself (sentinel-filled) is polled with workgroup-scope
//    atomics.
// Results always go to x.  Segments are taken by ticket, largest first.
constexpr int kLocalBlock = 1024;
constexpr int kLocalGroup = 8;  // dependencies polled together
constexpr int kLocalPre = 8;    // rows of prefetch per wavefront
constexpr int kLocalStaticLds = 20 * 1024;  // k_local's static LDS (staging, frontier, ticket), rounded up

// segment {lo, hi, bandwidth, longest row's off-diagonal count}: ring mode
// needs the bandwidth plus two rounds of rows in flight within the ring and
// rows whose entries and diagonal fit one 64-lane load
__device__ __forceinline__ bool local_ring_fits(int4 sg, int R)
{
    return sg.z + 2 * (kLocalBlock / kWave) <= R && sg.w < kWave;
}

// The rows of one segment for one wavefront of k_local, ring mode (rows of
// at most 64 off-diagonal entries).  The wavefront's rows m = 0, 1, ... (row
// lo + w + m W) go in batches of kLocalPre: at the start of a batch the
// registers prefetched during the previous batch (each row's first 64
// entries with the diagonal in lanes, its right-hand sides in lanes < k) are
// copied -- the only wait for global memory, on loads issued a whole batch
// earlier -- and only then are the next batch's loads issued, into other
// registers, so that solving the batch never waits on a fresh load.  Row
// bounds (lanes 0-7: rp[row], lanes 8-15: rp[row + 1]) run one batch ahead
// of the entries.  The watchdog only sets a flag, reported once at the end.
template <typename T, bool K1>
__device__ __forceinline__ void local_rows_ring(const int *__restrict__ rp, const int *__restrict__ ci,
                                                const T *__restrict__ val, const T *__restrict__ b,
                                                typename VBits<T>::U *__restrict__ xu, const SyncCtx &sc, int k,
                                                int4 sg, int R, typename VBits<T>::U *rv, int *rt, int *s_cur,
                                                typename VBits<T>::U *xs, T *vs)
{
    typedef typename VBits<T>::U U;
    constexpr U kS = VBits<T>::kSentinel;
    constexpr int kWaves = kLocalBlock / kWave;
    constexpr int B = kLocalPre;
    static_assert(2 * B <= kWave, "bounds of a batch fit one register");
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int qk = min(lane, k - 1);
    const int lo = sg.x, hi = sg.y, bw = sg.z;
    auto row_of = [&](int m) { return lo + w + m * kWaves; };
    auto poll = [&](int c) -> U {
        const int sl = c & (R - 1);
        const int tg = __hip_atomic_load(rt + sl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        asm volatile("" ::: "memory");  // tag before value (the writer stores the value first)
        const U v = __hip_atomic_load(rv + (size_t)sl * k + qk, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return tg == c ? v : kS;
    };
    auto load_bounds = [&](int mb) {
        const int r = row_of(mb + (lane & (B - 1)));
        return (lane < 2 * B && r < hi) ? rp[r + (lane >= B)] : 0;
    };
    // loads of batch mb from its bounds register: entries into c/v, right-hand
    // sides into bq; row bounds kept as scalars in a/e (e = diagonal's index)
    int nc[B], na[B], ne[B];
    T nv[B], nbq[B];
    auto issue_batch = [&](int mb, int bnd) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int r = row_of(mb + u);
            na[u] = __builtin_amdgcn_readlane(bnd, u);
            ne[u] = __builtin_amdgcn_readlane(bnd, u + B) - 1;
            if (r < hi) {
                if (lane <= ne[u] - na[u]) {
                    nc[u] = ci[na[u] + lane];
                    nv[u] = val[na[u] + lane];
                }
                nbq[u] = b[(long long)r * k + qk];
            }
        }
    };
#pragma unroll
    for (int u = 0; u < B; ++u) {
        nc[u] = 0;
        nv[u] = (T)0;
        nbq[u] = (T)0;
    }
    int bnd_next = load_bounds(0);
    {
        const int bnd = bnd_next;  // waits once, at the start
        issue_batch(0, bnd);
        bnd_next = load_bounds(B);
    }
    int front = lo;
    bool failed = false;
    for (int mb = 0; row_of(mb) < hi; mb += B) {
        // the previous batch's loads (and this batch's bounds) are complete: copy
        int cc[B], ca[B], ce[B];
        T cv[B], cbq[B];
#pragma unroll
        for (int u = 0; u < B; ++u) {
            cc[u] = nc[u];
            cv[u] = nv[u];
            cbq[u] = nbq[u];
            ca[u] = na[u];
            ce[u] = ne[u];
        }
        const int bnd = bnd_next;
        if (row_of(mb + B) < hi) {
            issue_batch(mb + B, bnd);
            bnd_next = load_bounds(mb + 2 * B);
        }
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int row = row_of(mb + u);
            if (row < hi) {
                const int a = ca[u], d = ce[u];
                const int cnt = d - a;
                Idle idle;
                T s = (T)0;
                if constexpr (K1) {
                    // one right-hand side: lane t polls entry t; values that
                    // arrive go to this wavefront's staging array, and the
                    // ready prefix is added in stored order from there (LDS
                    // reads at a wave-uniform address: a broadcast)
                    if (lane < cnt) vs[lane] = cv[u];
                    U xt = kS;
                    bool rdy = lane >= cnt;
                    int used = 0;
                    while (used < cnt) {
                        if (!rdy) {
                            xt = poll(cc[u]);
                            rdy = xt != kS;
                            if (rdy) xs[lane] = xt;
                        }
                        const unsigned long long ready = __ballot(rdy);
                        const int upto = min(cnt, ~ready ? (int)__builtin_ctzll(~ready) : kWave);
                        // a row with at most two entries left is about to be on
                        // the critical path: raise its issue priority early
                        // (measured 110 -> 104 ms on the deep band)
                        if (cnt - upto <= 2) __builtin_amdgcn_s_setprio(1);
                        if (upto == used) {
                            if (expired(sc, idle)) {
                                failed = true;
                                s = as_val(VBits<T>::kQNaN);
                                break;
                            }
                            continue;
                        }
                        // the rest of this row (chain tail, division, publish)
                        // is the critical path: take issue priority over the
                        // wavefronts that are still polling (reset after the
                        // publish); measured 117 -> 109 ms on the deep band
                        if (upto == cnt) __builtin_amdgcn_s_setprio(2);
                        idle.progress();
                        int t = used;
                        for (; t + 4 <= upto; t += 4) {
                            const T x0 = as_val(xs[t]), x1 = as_val(xs[t + 1]), x2 = as_val(xs[t + 2]),
                                    x3 = as_val(xs[t + 3]);
                            const T v0 = vs[t], v1 = vs[t + 1], v2 = vs[t + 2], v3 = vs[t + 3];
                            s = acc(sc, v0, x0, s);
                            s = acc(sc, v1, x1, s);
                            s = acc(sc, v2, x2, s);
                            s = acc(sc, v3, x3, s);
                        }
                        for (; t < upto; ++t) s = acc(sc, vs[t], as_val(xs[t]), s);
                        used = upto;
                    }
                } else
                for (int u0 = 0; u0 < cnt; u0 += kLocalGroup) {
                    int cu[kLocalGroup];
                    U xr[kLocalGroup];
#pragma unroll
                    for (int g = 0; g < kLocalGroup; ++g) {
                        cu[g] = __builtin_amdgcn_readlane(cc[u], min(u0 + g, cnt - 1));
                        xr[g] = u0 + g < cnt ? poll(cu[g]) : (U)0;
                    }
#pragma unroll
                    for (int g = 0; g < kLocalGroup; ++g) {
                        if (u0 + g < cnt) {
                            while (xr[g] == kS) {
                                if (expired(sc, idle)) {
                                    failed = true;
                                    xr[g] = VBits<T>::kQNaN;
                                    break;
                                }
                                xr[g] = poll(cu[g]);
                            }
                            s = acc(sc, bcast(cv[u], u0 + g), as_val(xr[g]), s);
                        }
                    }
                }
                __builtin_amdgcn_s_setprio(2);  // division and publish: the critical path
                // the slot's previous row and all its readers must be done
                idle.progress();
                while (row - R + bw >= front) {
                    int f = hi;
#pragma unroll
                    for (int v = 0; v < kWaves; ++v)
                        f = min(f, __hip_atomic_load(s_cur + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
                    front = f;
                    if (row - R + bw < front) break;
                    if (expired(sc, idle)) {
                        failed = true;
                        break;
                    }
                }
                if (lane < k) {
                    const long long o = (long long)row * k + lane;
                    U uo = as_bits((cbq[u] - s) / bcast(cv[u], cnt));
                    if (uo == kS) uo = VBits<T>::kQNaN;
                    const int sl = row & (R - 1);
                    __hip_atomic_store(rv + (size_t)sl * k + lane, uo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    asm volatile("" ::: "memory");  // value before tag
                    if (lane == 0) __hip_atomic_store(rt + sl, row, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    xu[o] = uo;
                }
                asm volatile("" ::: "memory");
                if (lane == 0) __hip_atomic_store(s_cur + w, row + kWaves, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __builtin_amdgcn_s_setprio(0);
            }
        }
    }
    if (__any(failed) && lane == 0) atomicOr(sc.err, 1);
}

// The rows of one segment for one wavefront of k_local, global mode (any
// row length): x itself is polled; each row's entries are loaded when the
// row starts.
template <typename T>
__device__ __forceinline__ void local_rows_global(const int *__restrict__ rp, const int *__restrict__ ci,
                                                  const T *__restrict__ val, const T *__restrict__ b,
                                                  typename VBits<T>::U *__restrict__ xu, const SyncCtx &sc, int k,
                                                  int4 sg, int *s_cur)
{
    typedef typename VBits<T>::U U;
    constexpr U kS = VBits<T>::kSentinel;
    constexpr int kWaves = kLocalBlock / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int qk = min(lane, k - 1);
    for (int row = sg.x + w; row < sg.y; row += kWaves) {
        const int a = rp[row], d = rp[row + 1] - 1;
        Idle idle;
        T s = (T)0;
        for (int j0 = a; j0 < d; j0 += kWave) {
            const int cnt = min(kWave, d - j0);
            const int cl = lane < cnt ? ci[j0 + lane] : 0;
            const T vl = lane < cnt ? val[j0 + lane] : (T)0;
            for (int u0 = 0; u0 < cnt; u0 += kLocalGroup) {
                int cu[kLocalGroup];
                U xr[kLocalGroup];
#pragma unroll
                for (int g = 0; g < kLocalGroup; ++g) {
                    cu[g] = __builtin_amdgcn_readlane(cl, min(u0 + g, cnt - 1));
                    xr[g] = u0 + g < cnt ? __hip_atomic_load(xu + (long long)cu[g] * k + qk, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_WORKGROUP)
                                         : (U)0;
                }
#pragma unroll
                for (int g = 0; g < kLocalGroup; ++g) {
                    if (u0 + g < cnt) {
                        while (xr[g] == kS) {
                            if (give_up(sc, idle)) {
                                xr[g] = VBits<T>::kQNaN;
                                break;
                            }
                            xr[g] = __hip_atomic_load(xu + (long long)cu[g] * k + qk, __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_WORKGROUP);
                        }
                        s = acc(sc, bcast(vl, u0 + g), as_val(xr[g]), s);
                    }
                }
            }
        }
        if (lane < k) {
            const long long o = (long long)row * k + lane;
            U uo = as_bits((b[o] - s) / val[d]);
            if (uo == kS) uo = VBits<T>::kQNaN;
            __hip_atomic_store(xu + o, uo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        asm volatile("" ::: "memory");
        if (lane == 0) __hip_atomic_store(s_cur + w, row + kWaves, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
}

template <typename T>
__global__ void __launch_bounds__(kLocalBlock)
    k_local(const int *__restrict__ rp, const int *__restrict__ ci, const T *__restrict__ val,
            const T *__restrict__ b, T *__restrict__ x, SyncCtx sc, int k, const int4 *__restrict__ seg, int n_seg,
            unsigned *__restrict__ ticket, int ring_log2)
{
    typedef typename VBits<T>::U U;
    constexpr int kWaves = kLocalBlock / kWave;
    extern __shared__ unsigned long long smem_local[];
    const int R = 1 << ring_log2;
    U *rv = reinterpret_cast<U *>(smem_local);              // R * k values
    int *rt = reinterpret_cast<int *>(rv + (size_t)R * k);  // R tags
    __shared__ int s_cur[kWaves];                            // row each wavefront is solving
    __shared__ U s_xs[kWaves][kWave];                        // k = 1: per-wavefront staging of x
    __shared__ T s_vs[kWaves][kWave];                        // and of the row's coefficients
    U *xu = reinterpret_cast<U *>(x);
    const int lane = threadIdx.x & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    ticket_loop(
        ticket, 1, n_seg,
        [&](int t) {
            if (t >= n_seg) return;
            const int4 sg = seg[t];
            if (local_ring_fits(sg, R))
                for (int i = threadIdx.x; i < R; i += blockDim.x) rt[i] = -1;
            if (lane == 0) s_cur[w] = sg.x + w;
        },
        [&](int t) {
            const int4 sg = seg[t];
            if (local_ring_fits(sg, R) && k == 1)
                local_rows_ring<T, true>(rp, ci, val, b, xu, sc, k, sg, R, rv, rt, s_cur, s_xs[w], s_vs[w]);
            else if (local_ring_fits(sg, R))
                local_rows_ring<T, false>(rp, ci, val, b, xu, sc, k, sg, R, rv, rt, s_cur, s_xs[w], s_vs[w]);
            else
                local_rows_global<T>(rp, ci, val, b, xu, sc, k, sg, s_cur);
            if (lane == 0) __hip_atomic_store(s_cur + w, sg.y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        });
}

// ------------------------------------------------ level steps in one workgroup
// CAP_ALGO_LEVEL.  One workgroup walks the matrix level by level: the rows
// of a level (in steps of at most B rows, B = the workgroup's 64 to
// kLevelBlock threads, chosen by the plan) are solved in parallel,
// one row per thread, and a workgroup barrier ends each step.  No thread
// waits on another row's value, so there is no polling and no readiness
// state; a step costs LDS reads, the row's chain and one barrier inside one
// compute unit, against the memory round trip of a publish -> observe hop
// between compute units.
//
// K right-hand sides (1, 2 or 4) are solved per row by the same thread,
// interleaved in b and x (row * K + h); the ring holds K values per slot.
//
// Values travel through an LDS ring, row i in slot i & mask.  The plan
// checks, for every off-diagonal entry c of a row in level L, that c > M_L -
// R (M_L: the largest row index in levels <= L; R = mask + 1): no row that
// maps to c's slot has been written by then.  Two slots follow the ring: R
// holds 0 (never written) and R + 1 takes the padding rows' results.
//
// The row data is laid out by the plan in step order (slot s * B +
// thread): {row, off-diagonal count}, the first E entries in stored order as
// ring slots (c & mask) and values, and the diagonal.  Unused entries are
// (slot R, value 0) and padding rows are {-1, 0} with diagonal 1, so every
// thread runs the same straight-line step: fma(0, 0, s) = s exactly (a sum
// that starts at +0 is never -0), and padding results go to slot R + 1 and
// to the sink.  Rows longer than E read the rest from the CSR (LONG).  The
// data is loaded D steps ahead into registers, the {row, count} pair 2D
// steps ahead (b is gathered through the row index); the barriers order LDS
// alone, so they do not wait for those loads.  The plan pads the steps to a
// multiple of 2D with padding slots.  Each sum is the thread-per-row
// kernels' stored-order chain and division: the same x bit for bit.
constexpr int kLevelBlock = 256;
constexpr int kLevelDepth4 = 10;  // steps of row data in flight with E = 4 entries carried
constexpr int kLevelDepth8 = 6;   // with E = 8 (the load counter holds 63 loads in flight)
constexpr int kLevelDepthK2 = 8;  // E = 4, 2 right-hand sides per row (registers)
constexpr int kLevelDepthK4 = 6;  // E = 4, 4 right-hand sides per row

template <typename T, int E, int K>
struct LevelStage {
    int row, cnt;
    int c[E];
    T v[E];
    T d, bo[K];
};

// 16-byte loads of E consecutive elements (E * sizeof(T) is a multiple of
// 16 for the instantiated E; the plan's arrays are 16-byte aligned)
template <typename V, int E>
__device__ __forceinline__ void load16(const V *__restrict__ src, V (&dst)[E])
{
    static_assert((E * sizeof(V)) % 16 == 0, "E values must fill 16-byte loads");
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
#pragma unroll
    for (int h = 0; h < (int)(E * sizeof(V) / 16); ++h) {
        const uint4 w = s4[h];
        __builtin_memcpy(reinterpret_cast<char *>(dst) + 16 * h, &w, 16);
    }
}

template <typename T, int E, int K>
__device__ __forceinline__ void level_load(LevelStage<T, E, K> &st, int2 rec, unsigned q, const int *__restrict__ lci,
                                           const T *__restrict__ lv, const T *__restrict__ ld, const T *__restrict__ b)
{
    st.row = rec.x;
    st.cnt = rec.y;
    load16<int, E>(lci + (size_t)q * E, st.c);
    load16<T, E>(lv + (size_t)q * E, st.v);
    st.d = ld[q];
    // padding slots read row 0's b, unused
    if constexpr (K == 1) {
        st.bo[0] = b[rec.x < 0 ? 0 : rec.x];
    } else {
        const size_t o = (size_t)(rec.x < 0 ? 0 : rec.x) * K;
#pragma unroll
        for (int h = 0; h < K; ++h) st.bo[h] = b[o + h];
    }
}

__device__ __forceinline__ void level_barrier()
{
    // only LDS carries values between steps: the fences order LDS alone, so
    // the barrier does not wait for the row-data loads in flight
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename T, int E, int D, bool UNFUSED, bool LONG, int K>
__global__ void __launch_bounds__(kLevelBlock)
    k_level(const int2 *__restrict__ rec, const int *__restrict__ lci, const T *__restrict__ lv,
            const T *__restrict__ ld, const int *__restrict__ rp, const int *__restrict__ ci,
            const T *__restrict__ val, const T *__restrict__ b, T *__restrict__ x, T *__restrict__ sink, int steps,
            int mask)
{
    const int B = blockDim.x;
    typedef typename VBits<T>::U U;
    extern __shared__ unsigned long long smem_level[];
    T *ring = reinterpret_cast<T *>(smem_level);
    const int t = threadIdx.x;
    LevelStage<T, E, K> st[D];
    int2 ahead[D], first[D];
    // the row pairs first, so no stage's b load waits behind another stage
#pragma unroll
    for (int r = 0; r < D; ++r) first[r] = rec[r * B + t];
#pragma unroll
    for (int r = 0; r < D; ++r) ahead[r] = rec[(D + r) * B + t];
#pragma unroll
    for (int r = 0; r < D; ++r) level_load<T, E, K>(st[r], first[r], r * B + t, lci, lv, ld, b);
    if (t < K) ring[(mask + 1) * K + t] = (T)0;
    level_barrier();
    // Two rounds of the D stages per iteration (steps is a multiple of 2D,
    // padded by the plan, so the body has no exit): the compiler's
    // load-counter waits are exact in the second round (the first round's,
    // merged at the loop head, assume about half the loads in flight).
    for (int s0 = 0; s0 < steps; s0 += 2 * D) {
#pragma unroll
        for (int h = 0; h < 2 * D; ++h) {
            const int r = h % D;
            const int s = s0 + h;
            const LevelStage<T, E, K> &w = st[r];
            T xv[E][K];
#pragma unroll
            for (int u = 0; u < E; ++u)
#pragma unroll
                for (int h = 0; h < K; ++h) xv[u][h] = ring[w.c[u] * K + h];
            T sum[K];
#pragma unroll
            for (int h = 0; h < K; ++h) {
                sum[h] = (T)0;
#pragma unroll
                for (int u = 0; u < E; ++u)
                    sum[h] = UNFUSED ? sum[h] + w.v[u] * xv[u][h] : fmac(w.v[u], xv[u][h], sum[h]);
            }
            if (LONG && w.cnt > E) {  // the rest of a long row, from the CSR
                const int j1 = rp[w.row] + w.cnt;
                for (int j = j1 - w.cnt + E; j < j1; ++j) {
                    const T a = val[j];
                    const int c = (ci[j] & mask) * K;
#pragma unroll
                    for (int h = 0; h < K; ++h) sum[h] = UNFUSED ? sum[h] + a * ring[c + h] : fmac(a, ring[c + h], sum[h]);
                }
            }
            const bool real = w.row >= 0;
            const int slot = (real ? w.row & mask : mask + 2) * K;
#pragma unroll
            for (int h = 0; h < K; ++h) {
                U o = as_bits((w.bo[h] - sum[h]) / w.d);
                if (o == VBits<T>::kSentinel) o = VBits<T>::kQNaN;  // as the publishing kernels do
                ring[slot + h] = as_val(o);
                if constexpr (K == 1)
                    *(real ? x + w.row : sink) = as_val(o);
                else
                    *(real ? x + (size_t)w.row * K + h : sink + h) = as_val(o);
            }
            // refill this stage for step s + D (padded: the arrays hold 2D
            // steps past the last), and its row pair for step s + 2D.  Kept
            // after the step's own work: interleaved with it, the compiler
            // parks in-flight load results in accumulation registers, and
            // each such move waits for its load.
            __builtin_amdgcn_sched_barrier(0);
            level_load<T, E, K>(st[r], ahead[r], (s + D) * B + t, lci, lv, ld, b);
            ahead[r] = rec[(s + 2 * D) * B + t];
            level_barrier();
        }
    }
}

// ------------------------------------------------ SyncFree CSC (K4 / K5)
// sptrsv_syncfree_cuda_analyser (CUDA/SyncFree_csc/spts_syncfree_cuda2.h:10-21)
__global__ void __launch_bounds__(kBlock) k_indegree(const int *__restrict__ idx, int nnz, int *__restrict__ deg)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < nnz) atomic_add_agent(deg + idx[g], 1);
}

// sptrsv_syncfree_cuda_executor_update (spts_syncfree_cuda2.h:24-86),
// redesigned for short columns: a group of G lanes per column (G chosen by
// the host from the mean column length; the reference uses one warp per
// column), kBlock / G columns per workgroup ticket.  A column waits until
// only its diagonal's own count remains, computes x_i = (b_i - left_sum_i) /
// diag, then its lanes scatter left_sum[r] += x_i * v and release one count
// of each row r.  Columns of one wavefront can depend on each other, so the
// loop is write-first with a wave-uniform exit (as in wrtfst_row).
// Ordering without cache-wide fences: the left-sum atomics return their old
// values and the in-degree decrements are issued only after those returned
// (a data dependency), so a reader that sees the count drop to 1 finds every
// contribution performed; it then reads left_sum with an agent-scope atomic
// load.  FORWARD takes the diagonal from the first entry of the column,
// BACKWARD from the last.
template <typename T, int G>
__global__ void __launch_bounds__(kBlock)
    k_csc(const int *__restrict__ cp, const int *__restrict__ ri, const T *__restrict__ cv,
          int *__restrict__ deg, T *__restrict__ left_sum, const T *__restrict__ b, T *__restrict__ x, int m,
          int backward, SyncCtx sc, unsigned *__restrict__ ticket)
{
    typedef typename VBits<T>::U U;
    constexpr int kCols = kBlock / G;
    const int g = threadIdx.x & (G - 1);
    ticket_loop(ticket, kCols, m, [&](int base) {
        int col = base + (int)(threadIdx.x / G);
        bool done = col >= m;
        if (!done && backward) col = m - 1 - col;
        int dpos = 0, start = 0, stop = 0;
        T bc = (T)0, dg = (T)1;  // loaded before the wait, off the critical path
        if (!done) {
            dpos = backward ? cp[col + 1] - 1 : cp[col];
            start = backward ? cp[col] : cp[col] + 1;
            stop = backward ? cp[col + 1] - 1 : cp[col + 1];
            bc = b[col];
            dg = cv[dpos];
        }
        Idle idle;
        do {  // wave-uniform exit
            if (!done) {
                const bool ready = ld_agent(deg + col) == 1;
                const bool failed = !ready && give_up(sc, idle);
                if (ready || failed) {
                    const T ls = as_val(ld_agent(reinterpret_cast<const U *>(left_sum) + col));
                    const T xi = failed ? as_val(VBits<T>::kQNaN) : (bc - ls) / dg;
                    T token = (T)0;
                    for (int j = start + g; j < stop; j += G)
                        token += __hip_atomic_fetch_add(left_sum + ri[j], xi * cv[j], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                    // the decrements depend on the returned sums: `one` passes
                    // through an asm that consumes token, so the compiler waits
                    // for the returning atomics before issuing the decrements
                    int one = 1;
                    asm volatile("" : "+v"(one) : "v"(token));
                    for (int j = start + g; j < stop; j += G) atomic_add_agent(deg + ri[j], -one);
                    if (g == 0) x[col] = xi;
                    asm volatile("" ::: "memory");
                    done = true;
                }
            }
        } while (__any(!done));
    });
}

// ---------------------------------------------------------------- adjoint
// grad_b = L^{-T} g: the CSC executor's BACKWARD mode
// (spts_syncfree_cuda2.h:50-55, :70-76) applied to the CSR of L read as the
// CSC of L^T, at thread granularity: one lane per node i (tickets taken in
// descending i), write-first so lanes of one wave never block each other.
// cnt[c] counts the off-diagonal entries of column c of L still unapplied.
// The same lane writes grad_val[j] = -grad_b[i] * x[col_j] for its row.
__global__ void __launch_bounds__(kBlock)
    k_adjoint(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
              const double *__restrict__ g, const double *__restrict__ xfwd, int *__restrict__ cnt,
              double *__restrict__ left_sum, double *__restrict__ grad_b, double *__restrict__ grad_val, int n,
              SyncCtx sc, unsigned *__restrict__ ticket)
{
    ticket_loop(ticket, kBlock, n, [&](int base) {
        const int t = base + threadIdx.x;
        if (t >= n) return;
        const int i = n - 1 - t;
        const int a = rp[i], d = rp[i + 1] - 1;
        const double gi = g[i], vd = val[d];  // before the wait, off the critical path
        Idle idle;
        bool done = false;
        do {  // wave-uniform exit, as in wrtfst_row
            if (!done) {
                const bool ready = ld_agent(cnt + i) == 0;
                const bool failed = !ready && give_up(sc, idle);
                if (ready || failed) {
                    // ordered as in k_csc: left_sum read with an agent-scope
                    // load after the count was seen; the count decrements
                    // depend on the returned left-sum atomics
                    const double ls = as_val(ld_agent(reinterpret_cast<const unsigned long long *>(left_sum) + i));
                    const double y = failed ? as_val(VBits<double>::kQNaN) : (gi - ls) / vd;
                    grad_b[i] = y;
                    double token = 0.0;
                    for (int j = a; j < d; ++j)
                        token += __hip_atomic_fetch_add(left_sum + ci[j], y * val[j], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
                    int one = 1;
                    asm volatile("" : "+v"(one) : "v"(token));
                    for (int j = a; j < d; ++j) atomic_add_agent(cnt + ci[j], -one);
                    if (grad_val)
                        for (int j = a; j <= d; ++j) grad_val[j] = -y * xfwd[ci[j]];
                    asm volatile("" ::: "memory");
                    done = true;
                }
            }
        } while (__any(!done));
    });
}

// grad_val[j] = -grad_b[row_j] * x[col_j] over the stored pattern (gather
// adjoint; the scatter adjoint fuses this into k_adjoint)
__global__ void __launch_bounds__(kBlock)
    k_grad_val(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ grad_b,
               const double *__restrict__ x, double *__restrict__ grad_val, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double y = grad_b[i];
    for (int j = rp[i]; j < rp[i + 1]; ++j) grad_val[j] = -y * x[ci[j]];
}

// off-diagonal column counts of L (the adjoint's in-degrees)
__global__ void __launch_bounds__(kBlock)
    k_offdiag_colcount(const int *__restrict__ rp, const int *__restrict__ ci, int n, int *__restrict__ cnt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int d = rp[i + 1] - 1;
    for (int j = rp[i]; j < d; ++j) atomic_add_agent(cnt + ci[j], 1);
}

// r = b - L x in FP64 (rows in CSR order, fused multiply-add chain), written
// as FP32 for the next correction solve; also max|x| for the convergence
// test.  Grid-stride over the n*k (row, right-hand side) pairs.
__global__ void __launch_bounds__(kBlock)
    k_residual32(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
                 const double *__restrict__ b, const double *__restrict__ x, float *__restrict__ r32, int n, int k,
                 unsigned long long *__restrict__ xmax_bits)
{
    const long long count = (long long)n * k;
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long m = 0;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < count; g += stride) {
        const int i = (int)(g / k), q = (int)(g - (long long)i * k);
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s = __builtin_fma(val[j], x[(long long)ci[j] * k + q], s);
        r32[g] = (float)(b[g] - s);
        const unsigned long long a = (unsigned long long)__double_as_longlong(fabs(x[g]));
        m = a > m ? a : m;
    }
    block_max_to(m, xmax_bits);
}

// r = b - L x in FP64 (one right-hand side, fma chain in CSR order) and
// max|x|: the residual of a refinement whose corrections are FP64 solves.
__global__ void __launch_bounds__(kBlock)
    k_residual64(const int *__restrict__ rp, const int *__restrict__ ci, const double *__restrict__ val,
                 const double *__restrict__ b, const double *__restrict__ x, double *__restrict__ r, int n, int k,
                 unsigned long long *__restrict__ xmax_bits)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long m = 0;
    // g = row * k + RHS: the k threads of a row read its entries together
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < (long long)n * k; g += stride) {
        const int i = (int)(g / k), q = (int)(g % k);
        double s = 0.0;
        for (int j = rp[i]; j < rp[i + 1]; ++j) s = __builtin_fma(val[j], x[(long long)ci[j] * k + q], s);
        r[g] = b[g] - s;
        const unsigned long long a = (unsigned long long)__double_as_longlong(fabs(x[g]));
        m = a > m ? a : m;
    }
    block_max_to(m, xmax_bits);
}

// x += d (FP64 correction) and max|d| (grid-stride)
__global__ void __launch_bounds__(kBlock)
    k_accumulate64(double *__restrict__ x, const double *__restrict__ d, long long count,
                   unsigned long long *__restrict__ dmax_bits)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long m = 0;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < count; g += stride) {
        x[g] += d[g];
        const unsigned long long a = (unsigned long long)__double_as_longlong(fabs(d[g]));
        m = a > m ? a : m;
    }
    block_max_to(m, dmax_bits);
}

// x += d (FP32 correction) and max|d| for the convergence test (grid-stride)
__global__ void __launch_bounds__(kBlock)
    k_accumulate(double *__restrict__ x, const float *__restrict__ dcor, long long count,
                 unsigned long long *__restrict__ dmax_bits, int first)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    unsigned long long m = 0;
    for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < count; g += stride) {
        const double dv = (double)dcor[g];
        x[g] = first ? dv : x[g] + dv;
        const unsigned long long a = (unsigned long long)__double_as_longlong(fabs(dv));
        m = a > m ? a : m;
    }
    block_max_to(m, dmax_bits);
}

__global__ void k_cast_f32(const double *__restrict__ a, float *__restrict__ o, long long count)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < count) o[g] = (float)a[g];
}

// dst row n-1-i = src row i (k interleaved values per row): an upper
// triangular plan solves the row- and column-reversed (lower) matrix
template <typename T>
__global__ void __launch_bounds__(kBlock) k_reverse_rows(const T *__restrict__ src, T *__restrict__ dst, int n, int k)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)n * k) return;
    const int i = (int)(g / k), q = (int)(g - (long long)i * k);
    dst[(long long)(n - 1 - i) * k + q] = src[g];
}

// Batch: local column indices -> columns of the block-diagonal matrix
// (matrix of each row found by binary search over row_off).
__global__ void __launch_bounds__(kBlock)
    k_globalize(const int *__restrict__ row_off, int n_mat, const int *__restrict__ rp,
                const int *__restrict__ ci_local, int n, int *__restrict__ ci)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = n_mat - 1;
    while (lo < hi) {  // largest m with row_off[m] <= i
        const int mid = (lo + hi + 1) >> 1;
        if (row_off[mid] <= i) lo = mid;
        else hi = mid - 1;
    }
    const int off = row_off[lo];
    for (int j = rp[i]; j < rp[i + 1]; ++j) ci[j] = ci_local[j] + off;
}

// ------------------------------------------------ partitioned banded solve
// For deep matrices whose lower bandwidth w = max(i - col) is at most the
// tail width W (the dependency chain is ~n/levels rows long and no row order
// can shorten it).  Rows are cut into P blocks of m >= W rows; block p's
// rows only reach back into the last W rows of block p-1.  With D the
// block-diagonal part of L, y = D^{-1} b (independent blocks) and V_p =
// D_pp^{-1} C_p (C_p: block p's entries in the coupling columns [pm - W,
// pm)), the exact solution is x_p = y_p - V_p t_{p-1} with t_p the last W
// values of x_p:
//   t_p = a_p + B_p t_{p-1},   a_p = tail(y_p),   B_p = -tail(V_p).
// The recurrence over the Q = P - 1 tails is scanned in two levels: in
// groups of gs consecutive tails sequentially (u_p from the group's first
// tail on), then over the group carries T_g in ceil(log2 G) rounds, then
// every tail is fixed up with its in-group product W_p = B_p ... B_first:
// t_p = u_p + W_p T_{g-1}.  The matrix products are formed once per plan
// (k_part_gemm); a solve only touches vectors.  The sums are reassociated,
// so results match the sequential solve to rounding, not bit for bit.
// The tail width W (the number of coupling columns, 128 or 256) is the
// smallest of those at least the matrix's lower bandwidth; blocks hold at
// least W rows.  Matrices are column-major W x W (element (c, k) at k * W +
// c); V comes from W / 64 batch solves of 64 interleaved columns (n x 64
// each, one after another).  The tables and the solve's work vectors are
// FP64 for FP32 plans too: b is widened when the block kernel reads it and x
// rounded once when the finish writes it.
constexpr int kPartW = 128;     // the narrower tail width
constexpr int kPartWMax = 256;  // the wider one

// Coupling right-hand sides (W / 64 zeroed slabs of 64 interleaved columns,
// slab h at r + h * n * 64): an entry of row i in block p >= 1 whose column c
// lies before the block goes to column q = c - (pm - W); duplicated entries
// add up.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_part_coupling(const int *__restrict__ rp, const int *__restrict__ ci, const T *__restrict__ val, int n, int m,
                    int W, double *__restrict__ r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int start = (i / m) * m;
    if (start == 0) return;
    const int base = start - W;
    for (int j = rp[i]; j < rp[i + 1]; ++j) {
        const int c = ci[j];
        const int q = c - base;
        if (c >= start || q < 0) continue;  // inside the block; q < 0 cannot occur (bandwidth checked on the host)
        r[(long long)(q >> 6) * n * 64 + (long long)i * 64 + (q & 63)] += (double)val[j];
    }
}

// V from the 64-column slabs the batch solves write (row-major, 64 per row)
// to column-major vt (column k at k * n): the finish then reads consecutive
// rows per column.
__global__ void __launch_bounds__(kBlock)
    k_part_transpose(const double *__restrict__ v, int n, int W, double *__restrict__ vt)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)n * W) return;
    const int k = (int)(g / n), i = (int)(g % n);
    vt[g] = v[(long long)(k >> 6) * n * 64 + (long long)i * 64 + (k & 63)];
}

// B_p = -(the last W rows of V_p) for p in [1, q_count), stored at index
// p - 1 (the first round's matrices).
__global__ void __launch_bounds__(kBlock)
    k_part_tail_matrix(const double *__restrict__ vt, int n, int m, int q_count, int W, double *__restrict__ b0)
{
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (long long)(q_count - 1) * W * W) return;
    const int c = (int)(g % W);
    const int k = (int)((g / W) % W);
    const int p = 1 + (int)(g / ((long long)W * W));
    const long long row = (long long)(p + 1) * m - W + c;
    b0[g] = -vt[(long long)k * n + row];
}

// y = D^{-1} b: one wavefront per block, rows in order.  A row's in-block
// entries in columns i-1 and i-2 are kept apart (near: their values, summed
// per column); the others (columns <= i-3) are stored ELL-style, ew <=
// kPartEll per row (value, and the byte offset of the column's slot in the
// LDS ring of the block's values: ((col - block start) mod 2W) * 8; pads
// point at slot 2W, which holds 0), the rest of a longer row in
// an overflow CSR (ov_rp per row, absolute columns).  Entry l of a row sits
// in lane l; its products with the block's earlier values are summed through
// LDS (8 lanes add 8 products each, then an 8-lane butterfly), and
//   y_i = ((b_i - S_i - w2_i y_{i-2}) - w1_i y_{i-1}) / d_i
// (as a multiplication by the reciprocal diagonal).  S_i needs only values
// up to y_{i-3}, so it is summed two rows ahead, off the chain from one row
// to the next.  Row data are loaded a group of kPartPre rows ahead.  The sum
// is reassociated and the division is a multiplication (tolerance, not
// bit-exact), like the whole partitioned solve.
constexpr int kPartEll = 64;
constexpr int kPartPre = 16;
// ring slots of a tail width W (rows' values by row mod 2W)
constexpr int part_ring(int W) { return 2 * W; }

// A double moved between lanes by DPP (the GFX9 DPP set).  Lanes whose
// source lies outside their row read 0 (bound_ctrl); lanes in rows outside
// ROW_MASK are left undefined (no old value to set up: two moves fewer per
// step).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The sum of v over each group of 8 lanes, in all 8 of them: butterflies by
// quad_perm [1,0,3,2], quad_perm [2,3,0,1] and row_half_mirror (lane i and
// 7 - i of each half row).  Each lane adds the same two halves, in either
// order, so all 8 hold the same bits.
__device__ __forceinline__ double part_sum8(double v)
{
    v += dpp_d<0xB1, 0xf>(v);
    v += dpp_d<0x4E, 0xf>(v);
    v += dpp_d<0x141, 0xf>(v);
    return v;
}

// v of lane `src` (0..63, may differ per lane; every lane active), through
// the LDS crossbar (ds_bpermute): no LDS memory and no VALU work
__device__ __forceinline__ double lane_read(double v, int src)
{
    const long long u = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src * 4, (int)u);
    const int hi = __builtin_amdgcn_ds_bpermute(src * 4, (int)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}


template <int W, int KR, typename T>
__global__ void __launch_bounds__(kWave)
    k_part_block(const int *__restrict__ ell_c, const double *__restrict__ ell_v, const int *__restrict__ ov_rp,
                 const int *__restrict__ ov_c, const double *__restrict__ ov_v, const double *__restrict__ inv_diag,
                 const double2 *__restrict__ near, const T *__restrict__ b, double *__restrict__ y,
                 double *__restrict__ tails, int n, int m, int ew, int k, int q_count)
{
    constexpr int kRing = part_ring(W);
    constexpr int kRow = KR + 3;  // a row's scalars in LDS: b per RHS, 1 / diagonal, near (i-1), near (i-2)
    __shared__ double ring[KR][kRing + 1];
    __shared__ double rowd[kPartPre][kRow];
    __shared__ double red[2 * KR][kWave];  // a row pair's lane products, transposed for the sums
    const int lane = threadIdx.x;
    // right-hand sides q0 .. q0 + KR - 1 (b and y interleaved, k per row)
    const long long q0 = (long long)blockIdx.y * KR;
    const int start = blockIdx.x * m;
    const int end = min(n, start + m);
    if (lane < KR) ring[lane][kRing] = 0.0;  // the pads' slot
    // a group's row data: entry `lane` of each of its kPartPre rows, and, in
    // lane r, row r's b values, reciprocal diagonal, near values and
    // overflow start (lane kPartPre: the end).  Lanes with nothing to load
    // are masked off (loading a clamped entry instead was slower).
    int nc[kPartPre];
    double nv[kPartPre], nb[KR], nd = 1.0, nw1 = 0.0, nw2 = 0.0;
    int no = 0;
    auto fetch = [&](int g0) {
        if (g0 + kPartPre <= end) {  // uniform: a whole group, one lane mask for all its loads
            if (lane < ew) {
#pragma unroll
                for (int r = 0; r < kPartPre; ++r) {
                    const long long e = (long long)(g0 + r) * ew + lane;
                    nc[r] = ell_c[e];
                    nv[r] = ell_v[e];
                }
            } else {
#pragma unroll
                for (int r = 0; r < kPartPre; ++r) {
                    nc[r] = kRing * 8;
                    nv[r] = 0.0;
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < kPartPre; ++r) {
                const long long e = (long long)(g0 + r) * ew + lane;
                const bool in = g0 + r < end && lane < ew;
                nc[r] = in ? ell_c[e] : kRing * 8;
                nv[r] = in ? ell_v[e] : 0.0;
            }
        }
        const int row = min(g0 + lane, end);
        const bool mine = lane < kPartPre && row < end;
#pragma unroll
        for (int j = 0; j < KR; ++j) nb[j] = mine ? (double)b[(long long)row * k + q0 + j] : 0.0;
        nd = mine ? inv_diag[row] : 1.0;
        const double2 w = mine ? near[row] : make_double2(0.0, 0.0);
        nw1 = w.x;
        nw2 = w.y;
        no = lane <= kPartPre ? ov_rp[row] : 0;
    };
    // The values go to y from the ring, a group at a time (one store of
    // kPartPre x KR values at the next group's top): a store per row would
    // make the waits for the prefetch (the memory counter covers stores too)
    // wait for it.
    static_assert(kPartPre * KR <= kWave, "a group's values are stored by one instruction");
    const int wr = lane / KR, wj = lane % KR;  // the (row, RHS) this lane stores
    auto store_group = [&](int s0) {
        const int i = s0 + wr;
        if (lane < kPartPre * KR && i < end) y[(long long)i * k + q0 + wj] = ring[wj][(i - start) & (kRing - 1)];
    };
    // Each lane carries the chain of one right-hand side, wj = lane mod KR:
    // y_{i-1}, y_{i-2} of that RHS (0 before the block's first row, where the
    // near values are 0 too).  Every lane still forms the products of its
    // ELL entry for all KR right-hand sides (the sums need them).
    double y1 = 0.0, y2 = 0.0;
    fetch(start);
    for (int g0 = start; g0 < end; g0 += kPartPre) {
        int cc[kPartPre];
        double cv[kPartPre];
#pragma unroll
        for (int r = 0; r < kPartPre; ++r) {
            cc[r] = nc[r];
            cv[r] = nv[r];
        }
        // the group's row scalars, read back as broadcasts row by row
        if (lane < kPartPre) {
#pragma unroll
            for (int j = 0; j < KR; ++j) rowd[lane][j] = nb[j];
            rowd[lane][KR] = nd;
            rowd[lane][KR + 1] = nw1;
            rowd[lane][KR + 2] = nw2;
        }
        const int og = no;
        // FAST: a whole group with no overflow entries (no branch between
        // its rows, so their sums and chains can interleave)
        const bool fast = g0 + kPartPre <= end &&
                          __builtin_amdgcn_readlane(og, 0) == __builtin_amdgcn_readlane(og, kPartPre);
        auto group = [&](auto fast_tag) {
            constexpr bool FAST = decltype(fast_tag)::value;
            // ring values of row r's ELL entries (all in columns the ring
            // holds; pads read the zero slot)
            auto ring_vals = [&](int r, double (&rv)[KR]) {
#pragma unroll
                for (int j = 0; j < KR; ++j) rv[j] = *(const double *)((const char *)ring[j] + cc[r]);
            };
            // the lane products of row r's S, from its ring values (plus its
            // overflow entries), into prod[0 .. KR)
            auto row_prod = [&](int r, const double (&rv)[KR], double *prod) {
#pragma unroll
                for (int j = 0; j < KR; ++j) prod[j] = cv[r] * rv[j];
                if constexpr (!FAST) {
                    const int o0 = __builtin_amdgcn_readlane(og, r), o1 = __builtin_amdgcn_readlane(og, r + 1);
                    for (int e = o0 + lane; e < o1; e += kWave) {  // rows longer than kPartEll
                        const int c = ov_c[e];
                        const double w = ov_v[e];
#pragma unroll
                        for (int j = 0; j < KR; ++j)
                            prod[j] = __builtin_fma(w, ring[j][(c - start) & (kRing - 1)], prod[j]);
                    }
                }
            };
            // S of rows r and r + 1 (all their right-hand sides): the 2 KR
            // sums of 64 lane products go through LDS, 8 lanes per sum add 8
            // products each (a fixed tree), and three butterfly steps join
            // the 8 lanes (part_sum8).  Every sum is formed the same way
            // whatever KR, so each right-hand side keeps its one-RHS bits.
            auto pair_sums = [&](int r, const double (&rva)[KR], const double (&rvb)[KR], double &sa, double &sb) {
                constexpr int N = 2 * KR;
                double prod[N];
                row_prod(r, rva, prod);
                row_prod(r + 1, rvb, prod + KR);
#pragma unroll
                for (int u = 0; u < N; ++u) red[u][lane] = prod[u];
                const double *src = &red[min(lane >> 3, N - 1)][(lane & 7) * 8];  // lanes past 8 N: unused
                double a[8];
#pragma unroll
                for (int t = 0; t < 8; ++t) a[t] = src[t];
                const double v = part_sum8(((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7])));
                // sum u sits in lanes 8u .. 8u + 7: this lane's RHS's two
                if constexpr (KR == 1) {
                    sa = bcast(v, 0);
                    sb = bcast(v, 8);
                } else {
                    sa = lane_read(v, 8 * wj);
                    sb = lane_read(v, 8 * (KR + wj));
                }
            };
            // the sums of the group's first two rows (columns below g0 - 1)
            double s0, s1, rv0[KR], rv1[KR];
            ring_vals(0, rv0);
            ring_vals(1, rv1);
            if (g0 + kPartPre < end) fetch(g0 + kPartPre);
            if (g0 > start) store_group(g0 - kPartPre);
            pair_sums(0, rv0, rv1, s0, s1);
            // y of row r from its sum s
            // y of row r for this lane's RHS from its sum s
            auto finish_row = [&](int r, double s) {
                const int i = g0 + r;
                const double di = rowd[r][KR], w1 = rowd[r][KR + 1], w2 = rowd[r][KR + 2];
                double t = rowd[r][wj] - s;
                t = __builtin_fma(-w2, y2, t);
                t = __builtin_fma(-w1, y1, t);
                const double yi = t * di;
                y2 = y1;
                y1 = yi;
                // every lane writes (the lanes of one RHS the same value to
                // the same slot), so no lane mask splits the rows' code
                ring[wj][(i - start) & (kRing - 1)] = yi;
            };
            // Rows in pairs: row r + 2's ring values (columns <= i - 1) are
            // read before y_i is written, row r + 3's (columns <= i) before
            // y_{i+1}; then both sums run, two independent DPP chains.
            static_assert(kPartPre % 2 == 0, "rows go in pairs");
#pragma unroll
            for (int r = 0; r < kPartPre; r += 2) {
                if constexpr (!FAST) {
                    if (g0 + r >= end) continue;  // uniform: the block's last group
                }
                double rv2[KR], rv3[KR];
                if (r + 2 < kPartPre) ring_vals(r + 2, rv2);
                finish_row(r, s0);
                if constexpr (!FAST) {
                    if (g0 + r + 1 >= end) continue;
                }
                if (r + 2 < kPartPre) ring_vals(r + 3, rv3);
                finish_row(r + 1, s1);
                if (r + 2 < kPartPre) pair_sums(r + 2, rv2, rv3, s0, s1);
            }
        };
        if (fast)
            group(std::true_type{});
        else
            group(std::false_type{});
    }
    store_group(start + (end - 1 - start) / kPartPre * kPartPre);  // the last group
    // a_p = the block's last W values, the tail scan's input (blocks p <
    // q_count, which are whole: m >= W rows), at (p * k + q) * W + c, from
    // the ring
    if ((int)blockIdx.x < q_count) {
#pragma unroll
        for (int c = lane; c < W; c += kWave) {
            const int slot = (end - W + c - start) & (kRing - 1);
#pragma unroll
            for (int j = 0; j < KR; ++j) tails[((long long)blockIdx.x * k + q0 + j) * W + c] = ring[j][slot];
        }
    }
}

// Plan-time matrix products C = A * B of the partition's W x W tables:
// ops[i] = {C, A, B} as matrix indices into bm (B < 0: C = A, a copy), one
// per blockIdx.x, each split into (W / 128)^2 quadrants of 128 x 128 by
// blockIdx.y.  Each of the 256 threads owns an 8 x 8 tile of its quadrant;
// the inner dimension is staged through LDS 16 at a time.
template <int W>
__global__ void __launch_bounds__(256) k_part_gemm(const int3 *__restrict__ ops, double *__restrict__ bm)
{
    constexpr long long kMat = (long long)W * W;
    constexpr int kQ = 128, kQn = W / kQ;
    const int3 op = ops[blockIdx.x];
    const double *A = bm + (long long)op.y * kMat;
    double *C = bm + (long long)op.x * kMat;
    const int t = threadIdx.x;
    const int qr = (int)blockIdx.y % kQn * kQ, qc = (int)blockIdx.y / kQn * kQ;  // quadrant: rows, columns
    if (op.z < 0) {  // copy (uniform) of the quadrant
        for (int e = t; e < kQ * kQ; e += 256) {
            const long long o = (long long)(qc + e / kQ) * W + qr + e % kQ;
            C[o] = A[o];
        }
        return;
    }
    const double *B = bm + (long long)op.z * kMat;
    __shared__ double sa[16][kQ];  // A(qr + c, j0 + jj)
    __shared__ double sb[16][kQ];  // B(j0 + jj, qc + k)
    const int c0 = (t & 15) * 8, k0 = (t >> 4) * 8;
    double acc[8][8];
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = 0.0;
    for (int j0 = 0; j0 < W; j0 += 16) {
        __syncthreads();
        for (int e = t; e < 16 * kQ; e += 256) {
            sa[e / kQ][e % kQ] = A[(long long)(j0 + e / kQ) * W + qr + e % kQ];
            sb[e % 16][e / 16] = B[(long long)(qc + e / 16) * W + j0 + e % 16];
        }
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < 16; ++jj) {
            double av[8], bv[8];
#pragma unroll
            for (int a = 0; a < 8; ++a) av[a] = sa[jj][c0 + a];
#pragma unroll
            for (int b = 0; b < 8; ++b) bv[b] = sb[jj][k0 + b];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int b = 0; b < 8; ++b) acc[a][b] = __builtin_fma(av[a], bv[b], acc[a][b]);
        }
    }
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) C[(long long)(qc + k0 + b) * W + qr + c0 + a] = acc[a][b];
}

// y_j = M v_j for one W x W column-major M and KQ vectors v_j in LDS:
// thread c returns row c of each (four partial sums, by index mod 4, in the
// same order for every KQ, so a right-hand side gets the same bits whichever
// chunk it is solved in).  M is read once for all KQ vectors.
template <int W, int KQ, bool NT>
__device__ __forceinline__ void part_matvec(const double *__restrict__ M, const double (*sv)[W], int c,
                                            double (&out)[KQ])
{
    double d[KQ][4];
#pragma unroll
    for (int j = 0; j < KQ; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u) d[j][u] = 0.0;
    constexpr int kUnroll = KQ >= 8 ? 2 : 8;
#pragma unroll kUnroll
    for (int kk = 0; kk < W; kk += 4) {
        double mv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            mv[u] = NT ? __builtin_nontemporal_load(M + (long long)(kk + u) * W + c) : M[(long long)(kk + u) * W + c];
#pragma unroll
        for (int j = 0; j < KQ; ++j)
#pragma unroll
            for (int u = 0; u < 4; ++u) d[j][u] = __builtin_fma(mv[u], sv[j][kk + u], d[j][u]);
    }
#pragma unroll
    for (int j = 0; j < KQ; ++j) out[j] = (d[j][0] + d[j][1]) + (d[j][2] + d[j][3]);
}

// The part of y_j = M v_j that columns k0 .. k0 + KLEN - 1 of M contribute,
// for row c (four partial sums by index mod 4, joined as in part_matvec).
template <int W, int KQ, int KLEN, bool NT>
__device__ __forceinline__ void part_matvec_cols(const double *__restrict__ M, const double (*sv)[W], int c, int k0,
                                                 double (&out)[KQ])
{
    double d[KQ][4];
#pragma unroll
    for (int j = 0; j < KQ; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u) d[j][u] = 0.0;
#pragma unroll 4
    for (int kk = k0; kk < k0 + KLEN; kk += 4) {
        double mv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            mv[u] = NT ? __builtin_nontemporal_load(M + (long long)(kk + u) * W + c) : M[(long long)(kk + u) * W + c];
#pragma unroll
        for (int j = 0; j < KQ; ++j)
#pragma unroll
            for (int u = 0; u < 4; ++u) d[j][u] = __builtin_fma(mv[u], sv[j][kk + u], d[j][u]);
    }
#pragma unroll
    for (int j = 0; j < KQ; ++j) out[j] = (d[j][0] + d[j][1]) + (d[j][2] + d[j][3]);
}

// The scan kernels that split a product's columns into kPartSplit quarters
// (one thread per row and quarter) and join the partial sums in LDS.
constexpr int kPartSplit = 4;

// The scan's vectors hold k right-hand sides: tail p of RHS q at (p * k + q)
// * W.  A workgroup takes the KQ right-hand sides from blockIdx.y * KQ;
// part_vec gives the j-th of them.
template <int W, int KQ>
__device__ __forceinline__ long long part_vec(int p, int k, int j)
{
    return ((long long)p * k + (long long)blockIdx.y * KQ + j) * W;
}

// Level 1 of the tail scan, one workgroup of kPartSplit x W threads per
// group of `gs` consecutive tails and chunk of KQ right-hand sides: u_first
// = a_first, u_p = a_p + B_p u_{p-1} (B_p at index p - 1 of bm).  The group's
// last u also goes to carry[g].  A step is sequential, so each product is
// split over the columns (thread: row c, quarter kq) to have four times the
// loads in flight; the first W threads own the vectors.  At W = 256 with
// several right-hand sides the workgroup is 16 wavefronts; holding it to 72
// VGPRs (7 waves per SIMD; 36 bytes per lane spill at KQ = 4) lets it share
// a CU with a 1024-thread workgroup of the thread-per-row solve (42 VGPRs)
// or of the fused refinement sweep (52), so a batch's partitioned solve runs
// beside the rest's persistent kernel instead of waiting for it.  Alone it
// is 2 % slower for that (758 against 743 us at 16 right-hand sides,
// n = 160,000).
template <int W, int KQ, bool NT>
__global__ void __launch_bounds__(kPartSplit * W, W >= 256 && KQ >= 2 ? 7 : 1)
    k_part_group_scan(const double *__restrict__ a, double *__restrict__ u, double *__restrict__ carry,
                      const double *__restrict__ bm, int gs, int q_count, int k)
{
    constexpr int kLen = W / kPartSplit;
    const int g = (int)blockIdx.x, t = (int)threadIdx.x, c = t % W, kq = t / W;
    const int p0 = g * gs, p1 = min(q_count, p0 + gs);
    __shared__ double sv[KQ][W];
    __shared__ double red[kPartSplit][KQ][W];
    double v[KQ];
    if (t < W) {
#pragma unroll
        for (int j = 0; j < KQ; ++j) {
            v[j] = a[part_vec<W, KQ>(p0, k, j) + c];
            u[part_vec<W, KQ>(p0, k, j) + c] = v[j];
        }
    }
    for (int p = p0 + 1; p < p1; ++p) {
        __syncthreads();  // the previous product is done with sv and red
        if (t < W) {
#pragma unroll
            for (int j = 0; j < KQ; ++j) sv[j][c] = v[j];
        }
        __syncthreads();
        double mv[KQ];
        part_matvec_cols<W, KQ, kLen, NT>(bm + (long long)(p - 1) * W * W, sv, c, kq * kLen, mv);
#pragma unroll
        for (int j = 0; j < KQ; ++j) red[kq][j][c] = mv[j];
        __syncthreads();
        if (t < W) {
#pragma unroll
            for (int j = 0; j < KQ; ++j) {
                v[j] = a[part_vec<W, KQ>(p, k, j) + c] + ((red[0][j][c] + red[1][j][c]) + (red[2][j][c] + red[3][j][c]));
                u[part_vec<W, KQ>(p, k, j) + c] = v[j];
            }
        }
    }
    if (t < W) {
#pragma unroll
        for (int j = 0; j < KQ; ++j) carry[part_vec<W, KQ>(g, k, j) + c] = v[j];
    }
}

// Level 3: t_p = u_p + W_p T_{g-1} for tails of groups g >= 1 (W_p at index
// p - gs of wm, T the scanned group carries), t_p = u_p in group 0.
template <int W, int KQ, bool NT>
__global__ void __launch_bounds__(W)
    k_part_fixup(const double *__restrict__ u, const double *__restrict__ tg, const double *__restrict__ wm, int gs,
                 double *__restrict__ t, int k)
{
    const int p = (int)blockIdx.x, c = (int)threadIdx.x;
    const int g = p / gs;
    __shared__ double sv[KQ][W];
    double v[KQ];
#pragma unroll
    for (int j = 0; j < KQ; ++j) v[j] = u[part_vec<W, KQ>(p, k, j) + c];
    if (g > 0) {  // uniform in the workgroup
#pragma unroll
        for (int j = 0; j < KQ; ++j) sv[j][c] = tg[part_vec<W, KQ>(g - 1, k, j) + c];
        __syncthreads();
        double mv[KQ];
        part_matvec<W, KQ, NT>(wm + (long long)(p - gs) * W * W, sv, c, mv);
#pragma unroll
        for (int j = 0; j < KQ; ++j) v[j] += mv[j];
    }
#pragma unroll
    for (int j = 0; j < KQ; ++j) t[part_vec<W, KQ>(p, k, j) + c] = v[j];
}


// One round of the log-depth scan over the group carries (kPartSplit
// workgroups of W threads per entry and chunk of KQ right-hand sides,
// each for a quarter of the rows: the round has only G entries, so each
// product is spread over four workgroups; thread: row, quarter of the
// columns, partial sums joined in LDS): dst_g = src_g + M_g src_{g-s} for g >= s (M_g at
// index g - s of the round's matrices), dst_g = src_g below s.
template <int W, int KQ, bool NT>
__global__ void __launch_bounds__(W)
    k_part_scan(const double *__restrict__ src, double *__restrict__ dst, const double *__restrict__ bm, int s, int k)
{
    constexpr int kRows = W / kPartSplit, kLen = W / kPartSplit;
    const int p = (int)blockIdx.x / kPartSplit, part = (int)blockIdx.x % kPartSplit;
    const int t = (int)threadIdx.x, r = t % kRows, kq = t / kRows, c = part * kRows + r;
    __shared__ double sv[KQ][W];
    __shared__ double red[kPartSplit][KQ][kRows];
    if (p >= s) {  // uniform in the workgroup
#pragma unroll
        for (int j = 0; j < KQ; ++j) sv[j][t] = src[part_vec<W, KQ>(p - s, k, j) + t];
        __syncthreads();
        double mv[KQ];
        part_matvec_cols<W, KQ, kLen, NT>(bm + (long long)(p - s) * W * W, sv, c, kq * kLen, mv);
#pragma unroll
        for (int j = 0; j < KQ; ++j) red[kq][j][r] = mv[j];
        __syncthreads();
        if (t < kRows) {
#pragma unroll
            for (int j = 0; j < KQ; ++j)
                dst[part_vec<W, KQ>(p, k, j) + c] = src[part_vec<W, KQ>(p, k, j) + c] +
                                                    ((red[0][j][r] + red[1][j][r]) + (red[2][j][r] + red[3][j][r]));
        }
    } else if (t < kRows) {
#pragma unroll
        for (int j = 0; j < KQ; ++j) dst[part_vec<W, KQ>(p, k, j) + c] = src[part_vec<W, KQ>(p, k, j) + c];
    }
}

// x = y - V t_{p-1} in blocks p >= 1 (t: the scan's result), x = y in
// block 0.  One thread per (row, RHS), g = row * k + q: V column-major, so
// for each column the lanes read consecutive rows (the k lanes of one row
// share an element).
template <int W, typename T>
__global__ void __launch_bounds__(kBlock)
    k_part_finish(const double *__restrict__ y, const double *__restrict__ vt, const double *__restrict__ t, int n,
                  int m, int k, T *__restrict__ x)
{
    const long long gi = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= (long long)n * k) return;
    const int i = k == 1 ? (int)gi : (int)(gi / k), q = k == 1 ? 0 : (int)(gi % k);
    const int p = i / m;
    double xi = y[gi];
    if (p > 0) {
        const double *tp = t + ((long long)(p - 1) * k + q) * W;
        double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll 8
        for (int c = 0; c < W; c += 4)
#pragma unroll
            for (int u = 0; u < 4; ++u)
                s[u] = __builtin_fma(__builtin_nontemporal_load(vt + (long long)(c + u) * n + i), tp[c + u], s[u]);
        xi -= (s[0] + s[1]) + (s[2] + s[3]);
    }
    x[gi] = (T)xi;
}

// The finish for chunks of KQ >= 2 right-hand sides: a workgroup takes
// W consecutive rows (at most two blocks, as m >= W) and RHS
// blockIdx.y * KQ + j.  It stages both blocks' t_{p-1} in LDS (RHS fastest,
// read as broadcasts), and each thread reads its row of V once for all KQ.
// The sums run in the single-RHS kernel's order (four partial sums by
// column mod 4), so each RHS gets that kernel's bits.  NT: V is read with
// non-temporal loads (one chunk covers every RHS, so V is read once; with
// several chunks, cached reads let later chunks reuse some of it).
template <int W, int KQ, typename T, bool NT>
__global__ void __launch_bounds__(W)
    k_part_finish_multi(const double *__restrict__ y, const double *__restrict__ vt, const double *__restrict__ t,
                        int n, int m, int k, T *__restrict__ x)
{
    __shared__ double st[2][W][KQ];
    const int lo = (int)blockIdx.x * W;
    const long long q0 = (long long)blockIdx.y * KQ;
    const int p_lo = lo / m, p_hi = (min(n, lo + W) - 1) / m;
    for (int e = (int)threadIdx.x; e < 2 * KQ * W; e += W) {
        const int sblk = e / (KQ * W), j = (e / W) % KQ, c = e % W;
        const int p = sblk == 0 ? p_lo : p_hi;
        st[sblk][c][j] = p > 0 ? t[((long long)(p - 1) * k + q0 + j) * W + c] : 0.0;
    }
    __syncthreads();
    const int i = lo + (int)threadIdx.x;
    if (i >= n) return;
    const int p = i / m;
    const int sblk = p == p_lo ? 0 : 1;
    double acc[KQ][4];
#pragma unroll
    for (int j = 0; j < KQ; ++j)
#pragma unroll
        for (int u = 0; u < 4; ++u) acc[j][u] = 0.0;
    if (p > 0) {
#pragma unroll 2
        for (int c = 0; c < W; c += 4) {
            double mv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                mv[u] = NT ? __builtin_nontemporal_load(vt + (long long)(c + u) * n + i) : vt[(long long)(c + u) * n + i];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int j = 0; j < KQ; ++j) acc[j][u] = __builtin_fma(mv[u], st[sblk][c + u][j], acc[j][u]);
        }
    }
#pragma unroll
    for (int j = 0; j < KQ; ++j) {
        const long long g = (long long)i * k + q0 + j;
        double xi = y[g];
        if (p > 0) xi -= (acc[j][0] + acc[j][1]) + (acc[j][2] + acc[j][3]);
        x[g] = (T)xi;
    }
}

}  // namespace cap
