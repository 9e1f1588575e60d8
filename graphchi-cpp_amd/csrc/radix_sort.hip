// Stable LSD radix sort of 64-bit keys (optional 32-bit payload), onesweep form.
//
// Replaces the serial PBBS sort of src/util/radixSort.hpp:49-116 (iSort with
// 8-bit digits, MAX_RADIX = 8).  Structure (Adinets & Merrill, "Onesweep",
// 2022):
//   1. one pass over the keys builds the digit histograms of every pass at once
//      (per-wave LDS histograms, then global atomics);
//   2. a tiny kernel turns each pass histogram into exclusive digit bases;
//   3. per digit pass, one kernel: a workgroup takes the next tile from an atomic
//      counter (so every earlier tile is already running), ranks its 4096 keys
//      stably with wavefront ballots (match-by-digit over 8 ballots), publishes
//      its per-digit counts, resolves the tile's global digit offsets by
//      decoupled look-back over the predecessors' published counts, reorders the
//      tile in LDS and writes digit runs out coalesced.
// Stability: keys are laid out wave-striped (wave w, item i, lane l holds key
// tile + w*1024 + i*64 + l) and ranked item by item in lane order, waves in
// order, so equal digits keep their input order -- the same order the
// reference's radixStep (counting sort, :50-66) produces.
#include "common.hpp"
#include "internal.hpp"

namespace gchi {

constexpr int kRsThreads = 256;
constexpr int kRsWaves = kRsThreads / kWave;
constexpr int kRsItems = 16;
constexpr int kRsTile = kRsThreads * kRsItems;  // 4096 keys
// Key-only passes use 512-thread workgroups (8192-key tiles): half as many
// tiles, so half as many tile tickets and look-back chains per pass.
constexpr int kRsThreadsK = 512;
constexpr int kRsTileK = kRsThreadsK * kRsItems;
constexpr int kRsRadix = 256;
constexpr int kRsMaxPasses = 8;
// Look-back status words: 2 flag bits (aggregate, inclusive) above a count.
// 32-bit words hold counts below 2^30 and serve every sort of fewer keys (a
// count never exceeds n); larger sorts use 64-bit words.  The 32-bit words
// halve the look-back reads and the per-pass clearing: 383 -> 356 us per pass
// at 67M keys.
template <typename SW>
struct StatusBits {
    static constexpr int kShift = 8 * (int)sizeof(SW) - 2;
    static constexpr SW kAgg = (SW)1 << kShift;
    static constexpr SW kInc = (SW)2 << kShift;
    static constexpr SW kVal = ((SW)1 << kShift) - 1;
};
constexpr uint32_t kSpinLimit = 1u << 26;  // default bound of one look-back wait
constexpr int kRsLookback = 4;
constexpr int kRsHistKeys = 8;  // keys per thread in flight in the histogram pass  // predecessor tiles read per look-back round trip

struct PassDesc {
    int shift[kRsMaxPasses];
    int bits[kRsMaxPasses];
};

// (src, dst) pairs read as the key (dst << id_bits) | src by a first pass, and
// written back as pairs by a last pass, so the pair sort needs no separate pack
// and unpack over HBM.  IO bit 1: the pass (or histogram) reads pairs; bit 2:
// the pass writes pairs.
struct PairIO {
    const uint32_t *src_in, *dst_in;
    uint32_t *src_out, *dst_out;
    int id_bits;
};
constexpr int kIoPackIn = 1, kIoUnpackOut = 2;

template <int IO>
GCHI_DEV uint64_t load_key(const uint64_t *kin, const PairIO &io, uint64_t i) {
    if (IO & kIoPackIn) return ((uint64_t)io.dst_in[i] << io.id_bits) | (uint64_t)io.src_in[i];
    return kin[i];
}

template <int IO>
__global__ __launch_bounds__(kRsThreads) void rs_histogram(const uint64_t *keys, uint64_t n, int npasses,
                                                           PassDesc pd, uint32_t *ghist, PairIO io) {
    __shared__ uint32_t s_h[kRsWaves][kRsMaxPasses][kRsRadix];  // 32 KB
    const int wave = threadIdx.x / kWave;
    for (int i = threadIdx.x; i < kRsWaves * kRsMaxPasses * kRsRadix; i += kRsThreads)
        (&s_h[0][0][0])[i] = 0;
    __syncthreads();
    // kRsHistKeys keys per thread are loaded before any is counted (one load in
    // flight per thread would leave the pass bound by memory latency).  Digit 0,
    // the common digit of the high passes for graph ids, is counted by one
    // ballot per wavefront instead of up to 64 LDS atomics on one word.
    const int lane = threadIdx.x & (kWave - 1);
    constexpr uint64_t kSpan = (uint64_t)kRsThreads * kRsHistKeys;
    for (uint64_t base = (uint64_t)blockIdx.x * kSpan; base < n; base += (uint64_t)gridDim.x * kSpan) {
        uint64_t kk[kRsHistKeys];
#pragma unroll
        for (int u = 0; u < kRsHistKeys; u++) {
            const uint64_t i = base + (uint64_t)u * kRsThreads + threadIdx.x;
            kk[u] = i < n ? load_key<IO>(keys, io, i) : 0ull;
        }
#pragma unroll
        for (int u = 0; u < kRsHistKeys; u++) {
            const bool valid = base + (uint64_t)u * kRsThreads + threadIdx.x < n;
            for (int p = 0; p < npasses; p++) {
                const uint32_t d = (uint32_t)(kk[u] >> pd.shift[p]) & ((1u << pd.bits[p]) - 1u);
                const uint64_t zero = __ballot(valid && d == 0u);
                if (valid && d != 0u) atomicAdd(&s_h[wave][p][d], 1u);
                if (lane == 0 && zero) atomicAdd(&s_h[wave][p][0], (uint32_t)__popcll(zero));
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < npasses * kRsRadix; i += kRsThreads) {
        int p = i / kRsRadix, d = i % kRsRadix;
        uint32_t c = 0;
        for (int w = 0; w < kRsWaves; w++) c += s_h[w][p][d];
        if (c) atomicAdd(&ghist[i], c);
    }
}

// ghist[p][d] -> exclusive digit bases, in place (one workgroup per pass).
__global__ void rs_hist_scan(uint32_t *ghist) {
    __shared__ uint32_t s[kRsRadix];
    uint32_t *h = ghist + blockIdx.x * kRsRadix;
    s[threadIdx.x] = h[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int d = 0; d < kRsRadix; d++) {
            uint32_t c = s[d];
            s[d] = run;
            run += c;
        }
    }
    __syncthreads();
    h[threadIdx.x] = s[threadIdx.x];
}

template <bool kVals, int T, int IO, typename SW>
__global__ __launch_bounds__(T) void rs_onesweep_pass(const uint64_t *__restrict__ kin,
                                                               const uint32_t *__restrict__ vin,
                                                               uint64_t *__restrict__ kout,
                                                               uint32_t *__restrict__ vout, uint64_t n, int shift,
                                                               int nbits, const uint32_t *__restrict__ gbase,
                                                               SW *status, uint32_t *tile_counter,
                                                               uint32_t *err, PairIO io, uint32_t spin_limit,
                                                               int force_stall) {
    constexpr int NW = T / kWave;
    constexpr int TILE = T * kRsItems;
    static_assert(T >= kRsRadix && T % kRsRadix == 0, "a thread per digit");
    __shared__ uint32_t s_wh[NW][kRsRadix];
    __shared__ uint32_t s_toff[kRsRadix];
    __shared__ uint32_t s_gb[kRsRadix];
    __shared__ uint64_t s_keys[TILE];
    __shared__ uint32_t s_vals[kVals ? TILE : 1];
    __shared__ uint32_t s_scan[kRsRadix / kWave];
    __shared__ uint32_t s_tile;

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wave = tid / kWave;
    const bool dig = tid < kRsRadix;  // threads 0..255 own digit tid (whole wavefronts)
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < NW * kRsRadix; i += T) (&s_wh[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint64_t tbase = (uint64_t)tile * TILE;
    const uint32_t tile_n = (uint32_t)min<uint64_t>(TILE, n - tbase);
    const uint32_t dmask = (1u << nbits) - 1u;

    uint64_t k[kRsItems];
    uint32_t v[kRsItems];
    uint32_t rank[kRsItems];
    const uint64_t wbase = tbase + (uint64_t)wave * (kRsItems * kWave);
#pragma unroll
    for (int i = 0; i < kRsItems; i++) {
        uint64_t idx = wbase + (uint64_t)i * kWave + lane;
        k[i] = idx < n ? load_key<IO>(kin, io, idx) : 0ull;
        if (kVals) v[i] = idx < n ? vin[idx] : 0u;
    }
    const uint64_t lt = lanemask_lt();
#pragma unroll
    for (int i = 0; i < kRsItems; i++) {
        uint64_t idx = wbase + (uint64_t)i * kWave + lane;
        bool valid = idx < n;
        uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < nbits; b++) {
            uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        uint32_t below = (uint32_t)__popcll(peers & lt);
        uint32_t cnt = (uint32_t)__popcll(peers);
        uint32_t b0 = 0;
        if (valid) b0 = s_wh[wave][d];
        __builtin_amdgcn_wave_barrier();
        if (valid && below == cnt - 1) s_wh[wave][d] = b0 + cnt;
        __builtin_amdgcn_wave_barrier();
        rank[i] = b0 + below;
    }
    __syncthreads();

    // per-digit: wave exclusive offsets and tile count (thread tid owns digit tid)
    uint32_t tcount = 0, incl = 0;
    typedef StatusBits<SW> B;
    SW *my_status = status + (size_t)tile * kRsRadix + (dig ? tid : 0);
    if (dig) {
        uint32_t run = 0;
#pragma unroll
        for (int w = 0; w < NW; w++) {
            uint32_t c = s_wh[w][tid];
            s_wh[w][tid] = run;
            run += c;
        }
        tcount = run;
        if (tile == 0)
            st_relaxed_agent(my_status, (SW)(B::kInc | (SW)tcount));
        else
            st_relaxed_agent(my_status, (SW)(B::kAgg | (SW)tcount));
        // tile-local exclusive scan over the 256 digits
        incl = tcount;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            uint32_t t = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += t;
        }
        if (lane == kWave - 1) s_scan[wave] = incl;
    }
    __syncthreads();
    if (dig) {
        uint32_t wb = 0;
        for (int w = 0; w < wave; w++) wb += s_scan[w];
        const uint32_t toff = wb + incl - tcount;
        s_toff[tid] = toff;

        // decoupled look-back for digit tid, kRsLookback predecessors per round
        // trip: their status words are loaded together, then consumed in order
        // (aggregates summed) up to the first inclusive one; a word not yet
        // published ends the round and is loaded again.  Tiles below 0 read as an
        // inclusive zero.
        uint64_t prefix = 0;
        if (tile > 0) {
            int64_t p = (int64_t)tile - 1;
            uint32_t spins = 0;
            for (;;) {
                SW sv[kRsLookback];
#pragma unroll
                for (int q = 0; q < kRsLookback; q++)
                    sv[q] = p - q >= 0 ? ld_relaxed_agent(status + (size_t)(p - q) * kRsRadix + tid) : B::kInc;
                int used = 0;
                bool inc = false, stalled = false;
#pragma unroll
                for (int q = 0; q < kRsLookback; q++) {
                    if (inc || stalled) continue;
                    const SW flag = force_stall ? (SW)0 : (SW)(sv[q] & ~B::kVal);
                    if (flag == 0) {
                        stalled = true;
                        continue;
                    }
                    prefix += (uint64_t)(sv[q] & B::kVal);
                    used++;
                    inc = flag == B::kInc;
                }
                if (inc) break;
                p -= used;
                if (used == 0) {
                    if (++spins > spin_limit) {
                        // the output is wrong from here on and the caller's
                        // status word says so; the scatter below drops any
                        // position outside [0, n)
                        atomicOr(err, GCHI_WORK_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            st_relaxed_agent(my_status, (SW)(B::kInc | ((SW)(prefix + tcount) & B::kVal)));
        }
        s_gb[tid] = gbase[tid] + (uint32_t)prefix - toff;
    }
    __syncthreads();

    // reorder the tile in LDS by digit
#pragma unroll
    for (int i = 0; i < kRsItems; i++) {
        uint64_t idx = wbase + (uint64_t)i * kWave + lane;
        if (idx < n) {
            uint32_t d = (uint32_t)(k[i] >> shift) & dmask;
            uint32_t pos = s_toff[d] + s_wh[wave][d] + rank[i];
            s_keys[pos] = k[i];
            if (kVals) s_vals[pos] = v[i];
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kRsItems; j++) {
        uint32_t p = (uint32_t)j * T + tid;
        if (p < tile_n) {
            uint64_t key = s_keys[p];
            uint32_t d = (uint32_t)(key >> shift) & dmask;
            uint32_t g = s_gb[d] + p;
            // In a correct sort g < n always holds.  After a look-back wait has
            // given up, the passes that follow read a buffer with unwritten
            // holes whose digits the histogram (taken from the input) never
            // counted, so their positions can run past n: those are dropped.
            if (g >= n) continue;
            if (IO & kIoUnpackOut) {
                io.src_out[g] = (uint32_t)(key & ((1ull << io.id_bits) - 1ull));
                io.dst_out[g] = (uint32_t)(key >> io.id_bits);
            } else {
                kout[g] = key;
            }
            if (kVals) vout[g] = s_vals[p];
        }
    }
}

__global__ void copy_u64(const uint64_t *a, uint64_t *b, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}
__global__ void copy_u32(const uint32_t *a, uint32_t *b, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}

static int npasses_for(int key_bits) { return (key_bits + 7) / 8; }

struct SortWork {
    uint32_t *err;  // status word (GCHI_WORK_* bits); first in the workspace
    uint64_t *ktmp;
    uint32_t *vtmp;
    uint32_t *ghist;
    uint64_t *status;  // one pass's look-back words (32- or 64-bit), cleared before every pass
    uint32_t *counters;  // [passes] tile counters
};

template <typename C>
static void sort_layout(C &c, uint64_t n, int key_bits, bool vals, SortWork *w) {
    int np = npasses_for(key_bits);
    uint64_t ntiles = div_up(n, kRsTile);
    auto *er = c.template take<uint32_t>(1);
    auto *kt = c.template take<uint64_t>(n);
    auto *vt = c.template take<uint32_t>(vals ? n : 1);
    auto *gh = c.template take<uint32_t>((size_t)np * kRsRadix);
    auto *st = c.template take<uint64_t>((size_t)ntiles * kRsRadix);
    auto *ct = c.template take<uint32_t>(np);
    if (w) *w = SortWork{er, kt, vt, gh, st, ct};
}

// Bound of every look-back wait and the stall switch of the test hook
// gchi_debug_sort_lookback (process-wide; read at each launch).
static uint32_t g_spin_limit = kSpinLimit;
static int g_force_stall = 0;

size_t sort_workspace_bytes(uint64_t n, int key_bits, bool vals) {
    SizerAdapter a;
    sort_layout(a, n, key_bits, vals, nullptr);
    return a.z.off;
}

// Sorts keys (and vals) on key bits [lo, hi), stably, in 8-bit digits from bit lo
// up.  The passes alternate between keys and a second buffer: alt when given,
// else one carved from the workspace.  With out == nullptr the result is copied
// back into keys; otherwise *out receives the buffer that holds it (keys or the
// second buffer) and nothing is copied.  Sorting on [0, b) and then on [b, hi)
// gives the same order as sorting on [0, hi) at once (every pass is stable).
// Key-only pass with the pair reads/writes chosen at run time.
template <typename SW>
static void launch_key_pass(int io_mode, const uint64_t *kin, uint64_t *kout, uint64_t n, int shift, int bits,
                            const uint32_t *gbase, SW *st, uint32_t *ctr, uint32_t *err, const PairIO &io,
                            uint64_t ntiles, hipStream_t s) {
    const uint32_t sl = g_spin_limit;
    const int fs = g_force_stall;
    const dim3 grid((unsigned)ntiles), blk(kRsThreadsK);
    const uint32_t *nv = nullptr;
    uint32_t *nvo = nullptr;
    switch (io_mode) {
    case 0:
        hipLaunchKernelGGL((rs_onesweep_pass<false, kRsThreadsK, 0, SW>), grid, blk, 0, s, kin, nv, kout, nvo, n,
                           shift, bits, gbase, st, ctr, err, io, sl, fs);
        break;
    case kIoPackIn:
        hipLaunchKernelGGL((rs_onesweep_pass<false, kRsThreadsK, kIoPackIn, SW>), grid, blk, 0, s, kin, nv, kout, nvo,
                           n, shift, bits, gbase, st, ctr, err, io, sl, fs);
        break;
    case kIoUnpackOut:
        hipLaunchKernelGGL((rs_onesweep_pass<false, kRsThreadsK, kIoUnpackOut, SW>), grid, blk, 0, s, kin, nv, kout,
                           nvo, n, shift, bits, gbase, st, ctr, err, io, sl, fs);
        break;
    default:
        hipLaunchKernelGGL((rs_onesweep_pass<false, kRsThreadsK, kIoPackIn | kIoUnpackOut, SW>), grid, blk, 0, s, kin,
                           nv, kout, nvo, n, shift, bits, gbase, st, ctr, err, io, sl, fs);
    }
}

// The sort itself; with pio, the keys come from pio's (src, dst) input and the
// result goes to its (src, dst) output (keys and alt are then the two
// intermediate buffers, and nothing is copied back).  A look-back wait that
// exceeds its bound sets GCHI_WORK_TIMEOUT in *err (the workspace's own status
// word when err is null, cleared here); a caller that passes its own word
// clears it and reads it back once the stream has run.
static int sort_impl(uint64_t *keys, uint64_t *alt, uint32_t *vals, uint64_t n, int lo, int hi, void *work,
                     size_t work_bytes, hipStream_t s, uint64_t **out, const PairIO *pio, uint32_t *err) {
    if (out) *out = keys;
    if ((n <= 1 || hi <= lo) && !work) return GCHI_OK;  // nothing to sort and no status word to clear
    GCHI_CHECK(lo >= 0 && hi <= 64, GCHI_ERR_INVALID, "sort: key bits outside [0, 64)");
    GCHI_CHECK(n < (1ull << 32), GCHI_ERR_INVALID, "sort: n must be < 2^32");
    GCHI_CHECK(!(out && vals), GCHI_ERR_INVALID, "sort: values are always copied back");
    GCHI_CHECK(n <= 1 || hi <= lo || keys, GCHI_ERR_INVALID, "sort: null keys");
    const bool hv = vals != nullptr;
    const int bits = hi > lo ? hi - lo : 0;
    GCHI_CHECK(work && work_bytes >= sort_workspace_bytes(n, bits, hv), GCHI_ERR_WORKSPACE,
               "sort: workspace too small");
    Carver c(work, work_bytes);
    SortWork w;
    sort_layout(c, n, bits, hv, &w);
    if (!err) {
        err = w.err;
        launch_fill_u32(err, 0u, 1, s);
    }
    if (n <= 1 || hi <= lo) return GCHI_OK;
    const int np = npasses_for(bits);
    const uint64_t ntiles = div_up(n, hv ? kRsTile : kRsTileK);  // <= the laid-out div_up(n, kRsTile)
    PassDesc pd;
    for (int p = 0; p < kRsMaxPasses; p++) {
        pd.shift[p] = lo + 8 * p;
        pd.bits[p] = p < np ? std::min(8, bits - 8 * p) : 0;
    }
    launch_fill_u32(w.ghist, 0u, (size_t)np * kRsRadix, s);
    launch_fill_u32(w.counters, 0u, np, s);
    unsigned hblocks = (unsigned)std::min<uint64_t>(ntiles, 1024);
    const PairIO io = pio ? *pio : PairIO{};
    if (pio)
        hipLaunchKernelGGL(rs_histogram<kIoPackIn>, dim3(hblocks), dim3(kRsThreads), 0, s, keys, n, np, pd, w.ghist,
                           io);
    else
        hipLaunchKernelGGL(rs_histogram<0>, dim3(hblocks), dim3(kRsThreads), 0, s, keys, n, np, pd, w.ghist, io);
    hipLaunchKernelGGL(rs_hist_scan, dim3(np), dim3(kRsRadix), 0, s, w.ghist);
    GCHI_LAUNCH_CHECK();
    uint64_t *ka = keys, *kb = alt ? alt : w.ktmp;
    uint32_t *va = vals, *vb = w.vtmp;
    const bool narrow = n < (1ull << 30);  // every count fits a 32-bit status word
    for (int p = 0; p < np; p++) {
        uint64_t *st = w.status;
        uint32_t *st32 = reinterpret_cast<uint32_t *>(w.status);
        const size_t nwords = (size_t)ntiles * kRsRadix;  // this pass's look-back words
        if (narrow) launch_fill_u32(st32, 0u, nwords, s);
        else launch_fill_u64(st, 0ull, nwords, s);
        const int iom = pio ? (p == 0 ? kIoPackIn : 0) | (p == np - 1 ? kIoUnpackOut : 0) : 0;
        if (hv && narrow)
            hipLaunchKernelGGL((rs_onesweep_pass<true, kRsThreads, 0, uint32_t>), dim3((unsigned)ntiles),
                               dim3(kRsThreads), 0, s, ka, va, kb, vb, n, pd.shift[p], pd.bits[p],
                               w.ghist + p * kRsRadix, st32, w.counters + p, err, PairIO{}, g_spin_limit,
                               g_force_stall);
        else if (hv)
            hipLaunchKernelGGL((rs_onesweep_pass<true, kRsThreads, 0, uint64_t>), dim3((unsigned)ntiles),
                               dim3(kRsThreads), 0, s, ka, va, kb, vb, n, pd.shift[p], pd.bits[p],
                               w.ghist + p * kRsRadix, st, w.counters + p, err, PairIO{}, g_spin_limit,
                               g_force_stall);
        else if (narrow)
            launch_key_pass(iom, ka, kb, n, pd.shift[p], pd.bits[p], w.ghist + p * kRsRadix, st32, w.counters + p,
                            err, io, ntiles, s);
        else
            launch_key_pass(iom, ka, kb, n, pd.shift[p], pd.bits[p], w.ghist + p * kRsRadix, st, w.counters + p, err,
                            io, ntiles, s);
        GCHI_LAUNCH_CHECK();
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (pio) return GCHI_OK;
    if (out) {
        *out = ka;
    } else if (ka != keys) {
        unsigned b = (unsigned)std::min<uint64_t>(div_up(n, 256), 65536);
        hipLaunchKernelGGL(copy_u64, dim3(b), dim3(256), 0, s, ka, keys, n);
        if (hv) hipLaunchKernelGGL(copy_u32, dim3(b), dim3(256), 0, s, va, vals, n);
        GCHI_LAUNCH_CHECK();
    }
    return GCHI_OK;
}

int radix_sort_u64_bits(uint64_t *keys, uint64_t *alt, uint32_t *vals, uint64_t n, int lo, int hi, void *work,
                        size_t work_bytes, hipStream_t s, uint64_t **out, uint32_t *err) {
    return sort_impl(keys, alt, vals, n, lo, hi, work, work_bytes, s, out, nullptr, err);
}

// Sorts keys (and vals) in place on their low key_bits bits.
int radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, void *work, size_t work_bytes,
                   hipStream_t s, uint32_t *err) {
    GCHI_CHECK(key_bits >= 0 && key_bits <= 64, GCHI_ERR_INVALID, "sort: key_bits outside [0, 64]");
    return radix_sort_u64_bits(keys, nullptr, vals, n, 0, key_bits, work, work_bytes, s, nullptr, err);
}

int read_work_status(const uint32_t *err, hipStream_t s) {
    uint32_t h = 0;
    GCHI_HIP(hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    GCHI_CHECK(!(h & GCHI_WORK_TIMEOUT), GCHI_ERR_TIMEOUT,
               "a bounded device wait (sort look-back) gave up: the output is invalid");
    GCHI_CHECK(!(h & GCHI_WORK_BAD_ID), GCHI_ERR_INVALID, "vertex id out of range");
    return GCHI_OK;
}

// ---- (src,dst) pair sort by (dst,src): sort_impl with PairIO ----

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_sort_workspace(uint64_t n, int key_bits, int with_values, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    *bytes = sort_workspace_bytes(n, key_bits, with_values != 0);
    return GCHI_OK;
}

extern "C" int gchi_sort_u64(uint64_t *d_keys, uint32_t *d_vals, uint64_t n, int key_bits, void *d_work,
                             size_t work_bytes, void *stream) {
    return radix_sort_u64(d_keys, d_vals, n, key_bits, d_work, work_bytes, as_stream(stream), nullptr);
}

// Key bits of the pair sort: at least 9, so that there are two passes and the
// pairs are read by one kernel and written by another (the output may be the
// input); the extra high bits are zero and leave the order as it is.
static int pair_key_bits(int id_bits) { return std::max(2 * id_bits, 9); }

extern "C" int gchi_sort_pairs_dst_src_workspace(uint64_t n, int id_bits, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    Sizer z;
    z.take<char>(sort_workspace_bytes(n, pair_key_bits(id_bits), false));  // status word first
    z.take<uint64_t>(n);
    *bytes = z.off;
    return GCHI_OK;
}

extern "C" int gchi_sort_pairs_dst_src(const uint32_t *d_src, const uint32_t *d_dst, uint64_t n, int id_bits,
                                       uint32_t *d_src_out, uint32_t *d_dst_out, void *d_work, size_t work_bytes,
                                       void *stream) {
    GCHI_CHECK(id_bits >= 1 && id_bits <= 32, GCHI_ERR_INVALID, "sort_pairs: id_bits must be in [1,32]");
    size_t need = 0;
    gchi_sort_pairs_dst_src_workspace(n, id_bits, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "sort_pairs: workspace too small");
    hipStream_t s = as_stream(stream);
    Carver c(d_work, work_bytes);
    size_t sw_bytes = sort_workspace_bytes(n, pair_key_bits(id_bits), false);
    void *sw = c.take<char>(sw_bytes);
    uint64_t *keys = c.take<uint64_t>(n);
    if (n == 0) return radix_sort_u64(nullptr, nullptr, 0, 0, sw, sw_bytes, s, nullptr);  // clears the status
    GCHI_CHECK(d_src && d_dst && d_src_out && d_dst_out, GCHI_ERR_INVALID, "sort_pairs: null pointer");
    const PairIO io{d_src, d_dst, d_src_out, d_dst_out, id_bits};
    if (n == 1) {  // nothing to sort: the pair is copied
        GCHI_TRY(radix_sort_u64(nullptr, nullptr, 0, 0, sw, sw_bytes, s, nullptr));
        if (d_src_out != d_src) GCHI_HIP(hipMemcpyAsync(d_src_out, d_src, 4, hipMemcpyDeviceToDevice, s));
        if (d_dst_out != d_dst) GCHI_HIP(hipMemcpyAsync(d_dst_out, d_dst, 4, hipMemcpyDeviceToDevice, s));
        return GCHI_OK;
    }
    return sort_impl(keys, nullptr, nullptr, n, 0, pair_key_bits(id_bits), sw, sw_bytes, s, nullptr, &io, nullptr);
}

extern "C" int gchi_work_status(const void *d_work, void *stream) {
    GCHI_CHECK(d_work, GCHI_ERR_INVALID, "work_status: null workspace");
    return read_work_status(static_cast<const uint32_t *>(d_work), as_stream(stream));
}

extern "C" int gchi_debug_sort_lookback(uint32_t spin_limit, int force_stall) {
    g_spin_limit = spin_limit ? spin_limit : kSpinLimit;
    g_force_stall = force_stall ? 1 : 0;
    return GCHI_OK;
}
