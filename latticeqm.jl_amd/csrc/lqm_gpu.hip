// liblqmgpu.so: host side of the gfx950 k-space band path behind the C ABI in
// include/lqm_gpu.h.  One context = one device + one stream + a grow-only workspace;
// one model = one Hops operator resident in HBM.  Work over a k-grid is cut into
// chunks sized to a workspace budget; each chunk runs
//   K_A k_hk_tridiag -> K_B k_tridiag_ql -> [K_C k_backtransform] -> [K_D phase contraction]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "lqm_gpu.h"
#include "aux_kernels.hpp"
#include "eig_kernels.hpp"
#include "wave_kernels.hpp"
#include "refine_kernels.hpp"
#include "hf_kernels.hpp"
#include "reg_kernels.hpp"
#include "col_kernels.hpp"
#include "wy_kernels.hpp"
#include "blocked_kernels.hpp"
#include "tri_kernels.hpp"
#include "response_kernels.hpp"
#include "partial_kernels.hpp"
#include "small_kernels.hpp"

using lqm::cx;

namespace {

thread_local std::string g_err;

lqm_status fail(lqm_status st, const std::string& msg) {
    g_err = msg;
    return st;
}

#define LQM_HIP(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(LQM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device-wide allocator calls (hipMalloc / hipFree) and the stream captures of SCF graphs are
// serialised by one mutex: a capture open on one thread's stream must not see another thread
// synchronise the device (the batch runs workers that allocate and free while others capture).
std::mutex g_cap_mu;
hipError_t dev_malloc(void** p, size_t bytes) {
    std::lock_guard<std::mutex> lk(g_cap_mu);
    return hipMalloc(p, bytes);
}
hipError_t dev_free(void* p) {
    if (!p) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_cap_mu);
    return hipFree(p);
}
template <typename T>
hipError_t dev_malloc(T** p, size_t bytes) {
    return dev_malloc(reinterpret_cast<void**>(p), bytes);
}

enum Slot {
    W_D, W_E, W_REFL, W_TAU, W_EVALS, W_PERM, W_ROT, W_SW, W_NSW, W_GA, W_GV, W_RHOK, W_EB,
    W_PART, W_SMALL, W_BANDS, W_KS, W_KW, W_OM, W_OUT, W_RR, W_Z, W_M, W_EV32, W_X, W_HA, W_W,
    W_S, W_G, W_LAM, W_STAT, W_GRADT, W_PV, W_PO, W_PW, W_KIDX, W_HK, W_KIDX2, W_BA, W_BV, W_BZ,
    W_BP, W_BVW, W_BT, W_BWT, W_BSC, W_BVEC, W_WYT, W_ZFMT, W_KLIST, W_TNORM, W_PHT, W_PHC, W_FOCC, W_MUE, W_X32, W_BDGB, W_BDGR, W_HFG, W_HFA, W_MA32, W_MX32, W_MUG, W_SCF, W_LNK, W_LNKO, W_RB, W_ROCC, W_RA, W_RC, W_RP, W_SPB, W_SPPH, W_SPM, W_SPT, W_SPI, W_SPC, W_SPKS, W_SMV, W_RRNEG, W_OASCAN, W_OAHDR, W_NSLOTS
};

constexpr size_t kWorkspaceBudget = size_t(12) << 30;  // bytes of per-k scratch per chunk
// per-k tridiagonal eigenpairs: a k-point whose eigenvalues come closer than this x ||T||
// (Gershgorin bound) takes the QL path, whose vectors stay orthogonal in clusters
constexpr double kTriClusterRel = 1e-3;
// blocks of one phase contraction launch (element tiles x k-splits x key groups): one full
// wave of three 256-thread blocks per CU on 256 CUs
constexpr int kContractBlocks = 768;

}  // namespace

// Device scratch of the allocation-free entry point lqm_dos_from_bands_dev_noalloc: the
// S x Nw DOS partials, at most 1024 x 256 doubles (2 MiB) for any Nw <= kScratchDoubles.
// Each device holds its own copy for the life of the process.  In host code the name is a
// host-side stand-in: kernels receive the device address that hipGetSymbolAddress returns,
// resolved once per device (scratch_address) when a context is created on it.
constexpr size_t kScratchDoubles = size_t(1024) * 256;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Address of this device's scratch array; call with `device` current.
lqm_status scratch_address(int device, double** out) {
    if (device < 0 || device >= kMaxDevices) return fail(LQM_ERR_UNSUPPORTED, "device ordinal above 63");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(LQM_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[device] = static_cast<double*>(ptr);
    }
    *out = g_scratch_ptr[device];
    return LQM_OK;
}
}  // namespace

// Kernel stages timed with HIP events when timing is on.
enum Stage {
    ST_TRIDIAG, ST_QL, ST_REPLAY, ST_BACKTRANSFORM, ST_ZFZ, ST_RHO2, ST_CONTRACT, ST_FOLD, ST_DOS,
    ST_REDUCE, ST_HK64, ST_GEMM_NN, ST_GEMM_CN, ST_GEMM_NC, ST_OA, ST_HF, ST_MIX, ST_MU, ST_BHK,
    ST_BTRD, ST_BUPD, ST_BEIG, ST_BVEC, ST_BBT, ST_BRHO, ST_OA2, ST_RESP, ST_SPF, ST_N
};
static_assert(ST_N == LQM_NSTAGES, "stage list and LQM_NSTAGES disagree");
const char* const kStageNames[ST_N] = {"hk_tridiag", "tridiag_ql", "replay", "backtransform", "zfz",
                                       "rho_twosided", "phase_contract", "phase_fold", "dos", "reduce",
                                       "herm_hk_fp64", "zgemm_nn_fp64", "zgemm_cn_fp64",
                                       "zgemm_nc_fp64", "oa_correct", "hf_meanfield", "mix",
                                       "mu_bisect", "blocked_hk", "blocked_panel", "blocked_update",
                                       "blocked_bisect", "blocked_invit", "blocked_backtransform",
                                       "blocked_rho", "oa_correct_later", "response_sums",
                                       "partial_filter"};

struct TimedLaunch {
    int stage;
    long long kpts;
    hipEvent_t a, b;
};

// Kernel variants the autotuner chooses between (per N).
enum VariantKind {
    VK_TRIDIAG = 0, VK_RHO = 1, VK_REFINE = 2, VK_QL = 3, VK_BT = 4, VK_BLOCKED = 5, VK_TRIEIG = 6, VK_N
};

struct lqm_context {
    std::map<int, int> tuned;  // (kind << 20 | N) -> variant
    lqm_progress_fn progress = nullptr;  // called as each model of a batch finishes
    void* progress_user = nullptr;
    int forced[VK_N] = {-1, -1, -1, -1, -1, -1, -1};
    int timing = 0;
    int rr_pairs = 0;  // the rho keys of the current pass are closed under R -> -R
    std::vector<TimedLaunch> pending;
    double stage_ms[ST_N] = {};
    long long stage_n[ST_N] = {};
    long long stage_k[ST_N] = {};
    int device = 0;
    int precision = 0;
    int lds_max = 65536;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    // side stream for the compact-WY T factors, which need only the reflectors and so run
    // beside the tridiagonal eigensolver (fork after the tridiagonalization, join before the
    // back-transformation); tfac_pending: a fork not joined yet
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool tfac_pending = false;
    void* buf[W_NSLOTS] = {};
    size_t cap[W_NSLOTS] = {};
    int* d_status = nullptr;
    void* pin = nullptr;     // pinned host bytes for the refinement's per-sweep header
    int refine_max = 3;      // Ogita-Aishima sweeps allowed per call (mixed precision)
    // upper bound on the per-k scratch of one chunk (bytes): the batch SCF gives each worker
    // context an equal share of the device's free memory so that workers on the largest
    // models cannot together exceed it
    size_t ws_cap = SIZE_MAX;
    double refine_thr = 1e-8;  // a k-point stops refining once its squared correction is below
    double report[6] = {};   // last mixed call: sweeps, max ||R||_F, max delta, near-cluster k,
                             // k-point sweeps run (sum over sweeps of the k-points refined),
                             // k-points whose rho_k the fused sweep formed
    const double* occ = nullptr;  // given occupations of the chunk in flight (device), or null
    int last_prec = -1;      // precision the last pass ran in (-1: none yet)
    double* scratch = nullptr;  // this device's scratch array (device address)
    int graphs = 1;          // SCF iterations replayed from captured HIP graphs where eligible
    bool capturing = false;  // a stream capture is open: workspace must not grow
};

struct lqm_model {
    lqm_context* ctx = nullptr;
    int N = 0, d = 0, nR = 0;
    std::vector<int> R;
    int* dR = nullptr;
    cx<double>* dT = nullptr;
};

// A sparse Bloch Hamiltonian: one CSR for all keys, entry e of row i = (column, key, value).
struct lqm_sparse_model {
    lqm_context* ctx = nullptr;
    int N = 0, d = 0, nR = 0;
    long long nnz = 0;
    std::vector<int> R;
    int* dR = nullptr;
    long long* rowptr = nullptr;
    int2* colkey = nullptr;
    cx<double>* val = nullptr;
    double bound = 0.0;  // max_i sum_e |t_e| >= ||H(k)||_2 at every k
};

namespace {

lqm_status ws(lqm_context* c, Slot s, size_t bytes, void** out) {
    if (bytes == 0) bytes = 16;
    if (c->cap[s] < bytes && c->capturing)
        return fail(LQM_ERR_DEVICE, "workspace growth requested inside a graph capture");
    if (c->cap[s] < bytes) {
        if (c->buf[s]) {
            hipError_t e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess && c->side) e = hipStreamSynchronize(c->side);
            if (e != hipSuccess) return fail(LQM_ERR_DEVICE, hipGetErrorString(e));
            (void)dev_free(c->buf[s]);
            c->buf[s] = nullptr;
            c->cap[s] = 0;
        }
        hipError_t e = dev_malloc(&c->buf[s], bytes);
        if (e != hipSuccess)
            return fail(LQM_ERR_DEVICE, "workspace allocation of " + std::to_string(bytes) +
                                            " bytes failed: " + hipGetErrorString(e));
        c->cap[s] = bytes;
    }
    *out = c->buf[s];
    return LQM_OK;
}

template <typename T>
lqm_status wsT(lqm_context* c, Slot s, size_t count, T** out) {
    void* p = nullptr;
    lqm_status st = ws(c, s, count * sizeof(T), &p);
    *out = static_cast<T*>(p);
    return st;
}

// Clears bytes (a multiple of 4) of device memory with a kernel: passes that an SCF graph
// captures avoid memset nodes.
void zero_dev(lqm_context* c, void* p, size_t bytes) {
    const size_t n = bytes / 4;
    if (n == 0) return;
    const unsigned g = (unsigned)std::min<size_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(lqm::k_zero_words, dim3(g), dim3(256), 0, c->stream, n, static_cast<uint32_t*>(p));
}

struct Plan {
    int N = 0, nR = 0;
    bool ldsA = false, ldsC = false;
    size_t shA = 0, shC = 0, shB = 0;
    int laneB = 64;
    int cap_rot = 0, cap_sw = 0;
    int rch = 0;  // rotation-staging capacity of K_C
    int nm = 0;   // register row length of k_replay (0: replay inside K_C)
    int nqA = 0;  // rows per lane of the one-wave tridiagonalization (0: 256-thread kernel)
    bool regA = false;  // tridiagonalization on a materialised H(k) (N <= 64)
    int colA = 0;       // ... 0: register tiles, 1: column per lane (variant 3), 2: the same with
                        // FP32 on one wave of 64 rows above 32 orbitals (variant 4)
    bool wyC = false;   // compact-WY back-transformation possible (N <= 64, Z from the replay)
    int btv = 1;        // back-transformation variant: 1 compact WY, 0 reflector loop
    size_t shAw = 0;
    bool rho2 = false;  // rho_k by two-sided reflectors on the packed Z f Z^T
    bool tri2 = false;  // tridiagonal eigenpairs by bisection + twisted factorisation (QL for clusters)
    bool trib = false;  // eigenvalues only, by bisection (N <= 64)
    size_t shM = 0, shR2 = 0;
    int Kc = 0;   // k-points per chunk
};

Plan make_plan(const lqm_context* c, int N, int nR, int Nk, bool vectors, bool rho,
               size_t cR = sizeof(double)) {
    Plan p;
    p.N = N;
    p.nR = nR;
    const size_t cC = 2 * cR;  // complex element of the eigensolver's working precision
    p.shA = cC * ((size_t)N * (N + 1) + 3 * N + 6) + sizeof(cx<double>) * nR;  // A, v[2], p, scalars
    p.ldsA = p.shA <= (size_t)c->lds_max;
    if (!p.ldsA) p.shA = cC * (3 * N + 6) + sizeof(cx<double>) * nR;
    p.rch = std::max(lqm::kRotChunk, N);
    const size_t smallC = cR * 2 * p.rch + cC * 4 * N + 8 * N + 32 + sizeof(int) * lqm::kHdrChunk + 16;
    p.shC = 2 * cR * (size_t)N * (N + 1) + smallC;
    p.ldsC = p.shC <= (size_t)c->lds_max;
    if (!p.ldsC) p.shC = smallC;
    p.cap_rot = 3 * N * N + 64 * N;
    p.cap_sw = 30 * N + 8;
    const size_t NN = (size_t)N * N;
    size_t perk = cC * NN + cC * N + 2 * cR * N + cR * N + 4 * N + 8;
    if (vectors || rho) perk += 2 * cR * (size_t)p.cap_rot + 4 * (size_t)p.cap_sw + 4;
    if (!p.ldsA) perk += cC * NN;
    if ((vectors || rho) && !p.ldsC) perk += cC * NN;
    if (rho) perk += 16 * NN;
    p.nm = N <= 8 ? 8 : N <= 16 ? 16 : N <= 32 ? 32 : N <= 64 ? 64 : 0;
    if ((vectors || rho) && p.nm) perk += cR * NN;
    const size_t np = (size_t)N * (N + 1) / 2;
    p.shAw = cC * (np + 2 * N) + sizeof(cx<double>) * nR;
    // forced variant, else the tuned one, else the default (1 = compact WY for the
    // back-transformation when N > 16, 0 otherwise: the WY kernel works on a 64 x 64 tile;
    // the column-per-lane tridiagonalization for N <= 64 -- measured fastest at every N from
    // 8 to 64 in both precisions; for N <= 8 the two-sided rho_k, whose tiles fit such
    // matrices -- 18.5 vs 21.1 ms for the 105-iteration SCF at 4 orbitals / 3600 k)
    auto variant = [&](int kind) {
        if (c->forced[kind] >= 0) return c->forced[kind];
        auto it = c->tuned.find((kind << 20) | N);
        if (it != c->tuned.end()) return it->second;
        if (kind == VK_TRIDIAG) return N <= 64 ? 3 : 0;
        if (kind == VK_RHO) return N <= 8 ? 1 : 0;
        return (kind == VK_BT && N > 16) || kind == VK_TRIEIG ? 1 : 0;
    };
    // QL: k-points (lanes) per block, 64 >> variant; fewer lanes per wave put more
    // independent recurrences on each SIMD
    {
        const int vq = variant(VK_QL);
        int L = 64 >> std::max(0, std::min(3, vq < 0 ? 0 : vq));
        while (L > 1 && 2 * cR * (size_t)N * L > (size_t)c->lds_max) L >>= 1;
        p.laneB = L;
        p.shB = 2 * cR * (size_t)N * L;
    }
    // tridiagonalization: 0 = 256-thread LDS-tile kernel, 1 = one wave, packed tile,
    // 2 = register tile on a materialised H(k) (N <= 64), 3 / 4 = column per lane on a
    // materialised H(k) (N <= 64; 3: one wave up to 32 orbitals, two above, 4: FP32 on one
    // wave of 64 rows above 32 orbitals)
    const int vt = variant(VK_TRIDIAG);
    if (vt == 1 && N <= 128 && p.shAw <= (size_t)c->lds_max) p.nqA = N <= 64 ? 1 : 2;
    if (vt >= 2 && N <= 64) {
        p.regA = true;
        p.colA = vt - 2;
        perk += 16 * NN;
    }
    // back-transformation: 0 = reflector by reflector (LDS), 1 = compact WY blocks of 16 on
    // the matrix cores (default for N <= 64, needs Z from the replay kernel)
    p.wyC = p.nm && N <= 64 && lqm::wy_lds_bytes() <= (size_t)c->lds_max;
    p.btv = variant(VK_BT);
    p.shM = sizeof(double) * ((size_t)N * (N + 1) + 2 + N + 4);
    p.shR2 = sizeof(cx<double>) * (np + 2 * N);
    // rho_k: 0 = V = QZ then V f V^H, 1 = two-sided reflectors on Z f Z^T
    p.rho2 = variant(VK_RHO) == 1 && rho && p.nm && p.shM <= (size_t)c->lds_max &&
             p.shR2 <= (size_t)c->lds_max;
    if (p.rho2) perk += 8 * NN;
    // tridiagonal eigenpairs: 0 = QL + rotation replay, 1 = bisection + twisted
    // factorisation with the QL path for k-points holding an eigenvalue cluster (feeds the
    // compact-WY back-transformation only)
    p.tri2 = (vectors || rho) && variant(VK_TRIEIG) == 1 && p.wyC && p.btv == 1 && !p.rho2;
    p.trib = !vectors && !rho && variant(VK_TRIEIG) == 1 && N <= 64;
    perk += 8;
    size_t kc = std::min(kWorkspaceBudget, c->ws_cap) / perk;
    if (kc < 1) kc = 1;
    p.Kc = (int)std::min<size_t>(kc, (size_t)Nk);
    return p;
}

// Brackets one kernel launch with events on the context stream (timing mode only).
struct StageScope {
    lqm_context* c;
    TimedLaunch t{};
    // kpts: k-points the bracketed launch processes (0 for stages not per k-point)
    StageScope(lqm_context* ctx, int stage, long long kpts = 0) : c(ctx) {
        if (!c->timing) return;
        t.stage = stage;
        t.kpts = kpts;
        if (hipEventCreate(&t.a) != hipSuccess || hipEventCreate(&t.b) != hipSuccess) {
            c = nullptr;
            return;
        }
        (void)hipEventRecord(t.a, c->stream);
    }
    ~StageScope() {
        if (!c || !c->timing) return;
        (void)hipEventRecord(t.b, c->stream);
        c->pending.push_back(t);
    }
};

template <typename K>
void allow_lds(K kernel, size_t bytes) {
    if (bytes > 65536)
        (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// Eigenvalues (and optionally the QL log) of one chunk: K_A then K_B, in the
// working precision R (double: the plain path; float: the bulk of the mixed path).
// A = Hermitian(H(k)) in FP64, full storage, for Kc k-points: 32 x 32 tile pairs held
// across 8 k-points per block, so the hopping stack is read from L2 once per tile and
// k-group rather than once per k-point.
// For nR <= 32 a phase table and one thread per matrix element that keeps its hopping
// entries in registers across up to 128 k-points (k_hk_stream); else the tiled kernel.
lqm_status launch_hk_full(lqm_context* c, lqm_model* m, int Kc, const double* ks, cx<double>* A) {
    const int N = m->N, nR = m->nR;
    if (nR > 0 && nR <= 32) {
        // 9 and 25: the 3x3 and 5x5 key sets of the nearest-image models, taken exactly
        const int nrm = nR <= 8 ? 8 : nR <= 9 ? 9 : nR <= 16 ? 16 : nR <= 24 ? 24 : nR <= 25 ? 25 : nR <= 28 ? 28 : 32;
        cx<double>* ph;
        lqm_status st = wsT(c, W_PHT, (size_t)Kc * nrm, &ph);
        if (st) return st;
        const size_t nph = (size_t)Kc * nrm;
        hipLaunchKernelGGL(lqm::k_phase_table, dim3((unsigned)((nph + 255) / 256)), dim3(256), 0, c->stream, Kc,
                           m->d, ks, nullptr, 1.0, nR, nrm, m->dR, 1.0, ph);
        const size_t NN = (size_t)N * N;
        const int be = (int)((NN + 255) / 256);
        // k-points per block: up to 128 while the grid keeps about 2048 blocks
        const int kb = (int)std::max<long long>(8, std::min<long long>(128, (long long)Kc * be / 2048));
        const dim3 g(be, (Kc + kb - 1) / kb);
        switch (nrm) {
            case 8: hipLaunchKernelGGL(lqm::k_hk_stream<8>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
            case 9: hipLaunchKernelGGL(lqm::k_hk_stream<9>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
            case 25: hipLaunchKernelGGL(lqm::k_hk_stream<25>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
            case 16: hipLaunchKernelGGL(lqm::k_hk_stream<16>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
            case 24: hipLaunchKernelGGL(lqm::k_hk_stream<24>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
            case 28: hipLaunchKernelGGL(lqm::k_hk_stream<28>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
            default: hipLaunchKernelGGL(lqm::k_hk_stream<32>, g, dim3(256), 0, c->stream, N, nR, m->dT, Kc, kb, ph, A); break;
        }
        return LQM_OK;
    }
    const int nt = (N + 31) / 32;
    const size_t sh = sizeof(cx<double>) * lqm::kHkK * std::max(1, nR);
    allow_lds(lqm::k_hk_full, sh);
    hipLaunchKernelGGL(lqm::k_hk_full, dim3(nt * (nt + 1) / 2, (Kc + lqm::kHkK - 1) / lqm::kHkK), dim3(256), sh,
                       c->stream, N, m->d, nR, m->dR, m->dT, Kc, ks, A);
    return LQM_OK;
}

// Column-per-lane tridiagonalization of N <= 64: one wave of 16 or 32 rows, two waves of 32
// rows above 32 orbitals (variant 4 with FP32: one wave of 64 rows).  Row chunks of 4.  FP32
// reads the row vectors from LDS (its multiply-adds already bound it); FP64 takes them by DPP
// broadcast inside the multiply-adds (LDS bound otherwise).  Register budgets for 4 (FP32) /
// 3 (FP64) waves per SIMD.
template <typename R>
lqm_status launch_tridiag_col(lqm_context* c, int variant, int N, int Kc, const cx<double>* H, R* dd, R* ee,
                              cx<R>* refl, cx<R>* tau) {
    constexpr bool f64 = sizeof(R) == 8;
    constexpr int BC = f64 ? 3 : 0, WPE = f64 ? 3 : 4;
    auto kA = lqm::k_tridiag_c<R, 2, 32, 4, BC, WPE>;
    int threads = 128;
    if (N <= 16) {
        kA = lqm::k_tridiag_c<R, 1, 16, 4, BC, 4>;
        threads = 64;
    } else if (N <= 32) {
        kA = lqm::k_tridiag_c<R, 1, 32, 4, BC, WPE>;
        threads = 64;
    } else if constexpr (!f64) {
        if (variant == 4) {
            kA = lqm::k_tridiag_c<R, 1, 64, 4, 0, 2>;
            threads = 64;
        }
    }
    hipLaunchKernelGGL(kA, dim3(Kc), dim3(threads), 0, c->stream, N, Kc, H, dd, ee, refl, tau);
    return LQM_OK;
}

// Hpre: H(k) of the chunk already materialised in FP64 (full Hermitian), or null.
template <typename R>
lqm_status launch_eigvals(lqm_model* m, const Plan& P, int Kc, const double* ks, R* evals,
                          bool logq, const cx<double>* Hpre = nullptr) {
    lqm_context* c = m->ctx;
    const int N = m->N;
    if (c->tfac_pending) {
        // a fork whose back-transformation never ran (an error return in between): the
        // reflectors are about to be overwritten, so the main stream waits for it here
        LQM_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
        c->tfac_pending = false;
    }
    R *dd, *ee;
    cx<R>*refl, *tau, *gA = nullptr;
    int* perm = nullptr;
    R* rot = nullptr;
    int *sw = nullptr, *nsw = nullptr;
    lqm_status st;
    if ((st = wsT(c, W_D, (size_t)N * Kc, &dd))) return st;
    if ((st = wsT(c, W_E, (size_t)N * Kc, &ee))) return st;
    if ((st = wsT(c, W_REFL, (size_t)N * N * Kc, &refl))) return st;
    if ((st = wsT(c, W_TAU, (size_t)N * Kc, &tau))) return st;
    if ((st = wsT(c, W_PERM, (size_t)N * Kc, &perm))) return st;
    if (!P.ldsA && (st = wsT(c, W_GA, (size_t)N * N * Kc, &gA))) return st;
    if (logq) {
        if ((st = wsT(c, W_ROT, (size_t)2 * P.cap_rot * Kc, &rot))) return st;
        if ((st = wsT(c, W_SW, (size_t)P.cap_sw * Kc, &sw))) return st;
        if ((st = wsT(c, W_NSW, (size_t)Kc, &nsw))) return st;
    }
    if (P.regA && !Hpre) {
        cx<double>* Hk;
        if ((st = wsT(c, W_HK, (size_t)N * N * Kc, &Hk))) return st;
        StageScope sc(c, ST_HK64, Kc);
        if ((st = launch_hk_full(c, m, Kc, ks, Hk))) return st;
        Hpre = Hk;
    }
    {
        StageScope sc(c, ST_TRIDIAG, Kc);
        if (P.regA && P.colA) {
            launch_tridiag_col<R>(c, P.colA + 2, N, Kc, Hpre, dd, ee, refl, tau);
        } else if (P.regA) {
            auto kA = lqm::k_tridiag_r<R>;
            hipLaunchKernelGGL(kA, dim3(Kc), dim3(256), 0, c->stream, N, Kc, Hpre, dd, ee, refl, tau);
        } else if (P.nqA == 1) {
            auto kA = lqm::k_hk_tridiag_w<R, 1>;
            allow_lds(kA, P.shAw);
            hipLaunchKernelGGL(kA, dim3(Kc), dim3(lqm::kWave), P.shAw, c->stream, N, m->d, m->nR, m->dR,
                               m->dT, ks, Kc, dd, ee, refl, tau);
        } else if (P.nqA == 2) {
            auto kA = lqm::k_hk_tridiag_w<R, 2>;
            allow_lds(kA, P.shAw);
            hipLaunchKernelGGL(kA, dim3(Kc), dim3(lqm::kWave), P.shAw, c->stream, N, m->d, m->nR, m->dR,
                               m->dT, ks, Kc, dd, ee, refl, tau);
        } else if (P.ldsA) {
            auto kA = lqm::k_hk_tridiag<R, true>;
            allow_lds(kA, P.shA);
            hipLaunchKernelGGL(kA, dim3(Kc), dim3(lqm::kTPB), P.shA, c->stream, N, m->d, m->nR, m->dR,
                               m->dT, ks, Kc, gA, dd, ee, refl, tau, Hpre);
        } else {
            auto kA = lqm::k_hk_tridiag<R, false>;
            allow_lds(kA, P.shA);
            hipLaunchKernelGGL(kA, dim3(Kc), dim3(lqm::kTPB), P.shA, c->stream, N, m->d, m->nR, m->dR,
                               m->dT, ks, Kc, gA, dd, ee, refl, tau, Hpre);
        }
    }
    LQM_HIP(hipGetLastError());
    const int L = P.laneB;
    const int gB = (Kc + L - 1) / L;
    int* klist = nullptr;
    if (logq && P.tri2) {
        // eigenpairs by bisection and twisted factorisations; the k-points whose cluster
        // vectors came out dependent are listed for the QL
        int* zfmt;
        R *Z, *tn;
        if ((st = wsT(c, W_ZFMT, (size_t)Kc, &zfmt))) return st;
        if ((st = wsT(c, W_KLIST, (size_t)Kc + 1, &klist))) return st;
        if ((st = wsT(c, W_Z, (size_t)N * N * Kc, &Z))) return st;
        if ((st = wsT(c, W_TNORM, (size_t)Kc, &tn))) return st;
        // the compact-WY T factors (reflectors only) on the side stream, beside the eigensolver;
        // below 2048 k-points the fork and join cost more than the overlap saves
        const int nbw = (N - 1 + lqm::kWyB - 1) / lqm::kWyB;
        if (nbw > 0 && Kc >= 2048) {
            cx<double>* Tg;
            if ((st = wsT(c, W_WYT, (size_t)nbw * lqm::kWyB * lqm::kWyB * Kc, &Tg))) return st;
            LQM_HIP(hipEventRecord(c->ev_fork, c->stream));
            LQM_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
            hipLaunchKernelGGL(lqm::k_wy_tfac<R>, dim3(nbw, Kc), dim3(64), 0, c->side, N,
                               static_cast<const cx<R>*>(c->buf[W_REFL]), static_cast<const cx<R>*>(c->buf[W_TAU]), Tg);
            LQM_HIP(hipGetLastError());
            LQM_HIP(hipEventRecord(c->ev_join, c->side));
            c->tfac_pending = true;
        }
        zero_dev(c, klist, sizeof(int));
        StageScope sc(c, ST_QL, Kc);
        const dim3 g4((Kc + 3) / 4);
        hipLaunchKernelGGL(lqm::k_tri_bisect<R>, g4, dim3(256), 0, c->stream, N, Kc, dd, ee, evals, perm, tn, c->d_status);
        switch (P.nm) {
            case 8: hipLaunchKernelGGL((lqm::k_tri_twist<R, 8>), g4, dim3(256), 0, c->stream, N, Kc, dd, ee, evals, tn, (R)kTriClusterRel, Z, zfmt, klist); break;
            case 16: hipLaunchKernelGGL((lqm::k_tri_twist<R, 16>), g4, dim3(256), 0, c->stream, N, Kc, dd, ee, evals, tn, (R)kTriClusterRel, Z, zfmt, klist); break;
            case 32: hipLaunchKernelGGL((lqm::k_tri_twist<R, 32>), g4, dim3(256), 0, c->stream, N, Kc, dd, ee, evals, tn, (R)kTriClusterRel, Z, zfmt, klist); break;
            default: hipLaunchKernelGGL((lqm::k_tri_twist<R, 64>), g4, dim3(256), 0, c->stream, N, Kc, dd, ee, evals, tn, (R)kTriClusterRel, Z, zfmt, klist); break;
        }
    }
    if (!logq && P.trib) {
        // eigenvalues only: bisection, no QL
        R* tn;
        if ((st = wsT(c, W_TNORM, (size_t)Kc, &tn))) return st;
        StageScope sc(c, ST_QL, Kc);
        hipLaunchKernelGGL(lqm::k_tri_bisect<R>, dim3((Kc + 3) / 4), dim3(256), 0, c->stream, N, Kc, dd, ee, evals,
                           perm, tn, c->d_status);
        LQM_HIP(hipGetLastError());
        return LQM_OK;
    }
    {
        StageScope sc(c, ST_QL, Kc);
        if (logq) {
            auto kB = lqm::k_tridiag_ql<R, true>;
            allow_lds(kB, P.shB);
            hipLaunchKernelGGL(kB, dim3(gB), dim3(L), P.shB, c->stream, N, Kc, dd, ee, evals, perm, rot,
                               sw, nsw, P.cap_rot, P.cap_sw, c->d_status, klist);
        } else {
            auto kB = lqm::k_tridiag_ql<R, false>;
            allow_lds(kB, P.shB);
            hipLaunchKernelGGL(kB, dim3(gB), dim3(L), P.shB, c->stream, N, Kc, dd, ee, evals, perm, rot,
                               sw, nsw, P.cap_rot, P.cap_sw, c->d_status, nullptr);
        }
    }
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

// zfmt (optional): k-points with zfmt[k] = 1 already hold their eigenvectors (k_tri_twist)
template <typename R, int NM>
void launch_replay(lqm_context* c, int N, int Kc, const R* rot, const int* sw, const int* nsw,
                   int cap_rot, int cap_sw, R* Z, const int* zfmt = nullptr) {
    StageScope sc(c, ST_REPLAY, Kc);
    auto kR = lqm::k_replay<R, NM>;
    hipLaunchKernelGGL(kR, dim3(Kc), dim3(lqm::kWave), 0, c->stream, N, rot, sw, nsw, cap_rot, cap_sw, Z, zfmt);
}

template <typename R, int MODE, bool ZIN>
void launch_kc(lqm_context* c, const Plan& P, int N, int Kc, const cx<R>* refl, const cx<R>* tau,
               const R* rot, const int* sw, const int* nsw, const R* evals, const int* perm,
               const double* kw, double wuni, double mu, double T, R* gV, const R* zin,
               cx<double>* out, double* eb) {
    StageScope sc(c, ST_BACKTRANSFORM, Kc);
    if (P.ldsC) {
        auto kC = lqm::k_backtransform<R, true, MODE, ZIN>;
        allow_lds(kC, P.shC);
        hipLaunchKernelGGL(kC, dim3(Kc), dim3(lqm::kTPB), P.shC, c->stream, N, Kc, refl, tau, rot, sw,
                           nsw, P.cap_rot, P.cap_sw, P.rch, evals, perm, kw, wuni, mu, T, c->occ, gV, zin,
                           out, eb);
    } else {
        auto kC = lqm::k_backtransform<R, false, MODE, ZIN>;
        allow_lds(kC, P.shC);
        hipLaunchKernelGGL(kC, dim3(Kc), dim3(lqm::kTPB), P.shC, c->stream, N, Kc, refl, tau, rot, sw,
                           nsw, P.cap_rot, P.cap_sw, P.rch, evals, perm, kw, wuni, mu, T, c->occ, gV, zin,
                           out, eb);
    }
}

template <typename R, int MODE>
lqm_status launch_backtransform(lqm_model* m, const Plan& P, int Kc, const R* evals,
                                const double* kw, double wuni, double mu, double T,
                                cx<double>* out, double* eb, cx<float>* out32 = nullptr,
                                bool* wrote32 = nullptr) {
    lqm_context* c = m->ctx;
    const int N = m->N;
    R* gV = nullptr;
    if (!P.ldsC) {
        lqm_status st = wsT(c, W_GV, (size_t)2 * N * N * Kc, &gV);
        if (st) return st;
    }
    auto* refl = static_cast<cx<R>*>(c->buf[W_REFL]);
    auto* tau = static_cast<cx<R>*>(c->buf[W_TAU]);
    auto* rot = static_cast<R*>(c->buf[W_ROT]);
    auto* sw = static_cast<int*>(c->buf[W_SW]);
    auto* nsw = static_cast<int*>(c->buf[W_NSW]);
    auto* perm = static_cast<int*>(c->buf[W_PERM]);
    if (P.nm) {
        R* Z;
        lqm_status st = wsT(c, W_Z, (size_t)N * N * Kc, &Z);
        if (st) return st;
        if (P.wyC && P.btv == 1) {
            const int* zfmt = P.tri2 ? static_cast<const int*>(c->buf[W_ZFMT]) : nullptr;
            switch (P.nm) {
                case 8: launch_replay<R, 8>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z, zfmt); break;
                case 16: launch_replay<R, 16>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z, zfmt); break;
                case 32: launch_replay<R, 32>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z, zfmt); break;
                default: launch_replay<R, 64>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z, zfmt); break;
            }
            LQM_HIP(hipGetLastError());
            StageScope sc(c, ST_BACKTRANSFORM, Kc);
            const int nbw = (N - 1 + lqm::kWyB - 1) / lqm::kWyB;
            cx<double>* Tg = nullptr;
            lqm_status st;
            if ((st = wsT(c, W_WYT, (size_t)std::max(1, nbw) * lqm::kWyB * lqm::kWyB * Kc, &Tg))) return st;
            if (c->tfac_pending) {
                // launched beside the eigensolver for this chunk: join the side stream
                LQM_HIP(hipStreamWaitEvent(c->stream, c->ev_join, 0));
                c->tfac_pending = false;
            } else if (nbw > 0) {
                hipLaunchKernelGGL(lqm::k_wy_tfac<R>, dim3(nbw, Kc), dim3(64), 0, c->stream, N, refl, tau, Tg);
            }
            if constexpr (std::is_same<R, float>::value && MODE == 0) {
                // mixed path: the FP32 vectors on the FP32 matrix cores, written in FP32 when
                // the caller takes them so (out32), else widened to FP64
                const size_t sh = lqm::wy32_lds_bytes();
                if (out32) {
                    allow_lds(lqm::k_backtransform_wy32<cx<float>>, sh);
                    hipLaunchKernelGGL(lqm::k_backtransform_wy32<cx<float>>, dim3(Kc), dim3(256), sh, c->stream, N,
                                       refl, Z, perm, out32, Tg, zfmt);
                    if (wrote32) *wrote32 = true;
                } else {
                    allow_lds(lqm::k_backtransform_wy32<cx<double>>, sh);
                    hipLaunchKernelGGL(lqm::k_backtransform_wy32<cx<double>>, dim3(Kc), dim3(256), sh, c->stream, N,
                                       refl, Z, perm, out, Tg, zfmt);
                }
            } else {
                auto kW = lqm::k_backtransform_wy<R, MODE>;
                const size_t sh = lqm::wy_lds_bytes();
                allow_lds(kW, sh);
                // MODE 1 (rho_k): a key list closed under R -> -R is contracted from the upper triangle
                hipLaunchKernelGGL(kW, dim3(Kc), dim3(256), sh, c->stream, N, refl, tau, Z, evals, perm, kw, wuni,
                                   mu, T, c->occ, out, eb, Tg, zfmt, (MODE == 1 && c->rr_pairs) ? 1 : 0);
            }
            LQM_HIP(hipGetLastError());
            return LQM_OK;
        }
        switch (P.nm) {
            case 8: launch_replay<R, 8>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
            case 16: launch_replay<R, 16>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
            case 32: launch_replay<R, 32>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
            default: launch_replay<R, 64>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
        }
        LQM_HIP(hipGetLastError());
        launch_kc<R, MODE, true>(c, P, N, Kc, refl, tau, rot, sw, nsw, evals, perm, kw, wuni, mu, T, gV, Z,
                                 out, eb);
    } else {
        launch_kc<R, MODE, false>(c, P, N, Kc, refl, tau, rot, sw, nsw, evals, perm, kw, wuni, mu, T, gV,
                                  nullptr, out, eb);
    }
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

// rho_k for N <= 64: replay -> M = Z f Z^T -> two-sided reflectors on packed M.
lqm_status launch_rho2(lqm_model* m, const Plan& P, int Kc, const double* evals, const double* kw,
                       double wuni, double mu, double T, cx<double>* rhok, double* eb) {
    lqm_context* c = m->ctx;
    const int N = m->N;
    auto* refl = static_cast<cx<double>*>(c->buf[W_REFL]);
    auto* tau = static_cast<cx<double>*>(c->buf[W_TAU]);
    auto* rot = static_cast<double*>(c->buf[W_ROT]);
    auto* sw = static_cast<int*>(c->buf[W_SW]);
    auto* nsw = static_cast<int*>(c->buf[W_NSW]);
    auto* perm = static_cast<int*>(c->buf[W_PERM]);
    double *Z, *M;
    lqm_status st;
    if ((st = wsT(c, W_Z, (size_t)N * N * Kc, &Z))) return st;
    if ((st = wsT(c, W_M, (size_t)N * N * Kc, &M))) return st;
    switch (P.nm) {
        case 8: launch_replay<double, 8>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
        case 16: launch_replay<double, 16>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
        case 32: launch_replay<double, 32>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
        default: launch_replay<double, 64>(c, N, Kc, rot, sw, nsw, P.cap_rot, P.cap_sw, Z); break;
    }
    LQM_HIP(hipGetLastError());
    {
        StageScope sc(c, ST_ZFZ, Kc);
        auto kM = lqm::k_zfz<double>;
        allow_lds(kM, P.shM);
        hipLaunchKernelGGL(kM, dim3(Kc), dim3(lqm::kTPB), P.shM, c->stream, N, Z, evals, perm, kw, wuni,
                           mu, T, c->occ, M, eb);
    }
    LQM_HIP(hipGetLastError());
    {
        StageScope sc(c, ST_RHO2, Kc);
        auto kR2 = lqm::k_rho_w<double, 1>;
        allow_lds(kR2, P.shR2);
        hipLaunchKernelGGL(kR2, dim3(Kc), dim3(lqm::kWave), P.shR2, c->stream, N, refl, tau, M, rhok);
    }
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

// k-splits of a contraction launch with `tiles` element tiles x key groups: kContractBlocks
// blocks in all (callers cap it by the chunk's k-points / kContractMinK, so that a small
// matrix is not cut into splits of a few k-points each, which the fold would then sum
// one partial at a time)
constexpr int kContractMinK = 32;
int contract_splits(int tiles) { return kContractBlocks / std::max(1, tiles); }

// Key groups of the phase contraction: RG keys per group, the smallest of these holding nRr.
int contract_group(int nRr) {
    // 9 and 25: the 3x3 and 5x5 key sets of the nearest-image models, taken exactly
    return nRr <= 1 ? 1 : nRr <= 4 ? 4 : nRr <= 8 ? 8 : nRr <= 9 ? 9 : nRr <= 16 ? 16 : nRr <= 24 ? 24
         : nRr <= 25 ? 25 : nRr <= 28 ? 28 : 32;
}

// rho(R) += sum_k w_k e^{-2 pi i k.R} rho_k over a chunk: the phase table of the chunk's
// k-points and keys (groups x RG columns), the split-K contraction, the fixed-order fold.
// With a key set closed under R -> -R (c->rr_pairs) the contraction forms the upper triangle
// only and the fold takes the lower one from rho(-R) (half the reads and multiply-adds).
lqm_status launch_contract(int RG, lqm_context* c, int N, int Kc, int d, const double* ks, const double* kw,
                           double wuni, int nRr, const int* Rr, const cx<double>* rhok, int ksplit, int S,
                           int groups, cx<double>* part, cx<double>* rho) {
    hipStream_t s = c->stream;
    const int NN = N * N;
    const int ldp = groups * RG;
    cx<double>* ph;
    lqm_status st = wsT(c, W_PHC, (size_t)Kc * ldp, &ph);
    if (st) return st;
    const bool herm = c->rr_pairs && nRr > 0;
    int* neg = nullptr;
    int ne = NN, Nh = 0;
    if (herm) {
        if ((st = wsT(c, W_RRNEG, (size_t)nRr, &neg))) return st;
        hipLaunchKernelGGL(lqm::k_neg_keys, dim3((nRr + 63) / 64), dim3(64), 0, s, nRr, d, Rr, neg);
        ne = N * (N + 1) / 2;
        Nh = N;
        // more k-splits over the smaller element range (the grid stays as wide), within part's
        // size (groups x S x RG x NN)
        const long long Smax = std::min<long long>((long long)S * NN / ne, Kc);
        const int tilesU = (ne + 255) / 256;
        const int Sw = std::max(1, std::min(std::max(1, Kc / kContractMinK), contract_splits(tilesU * groups)));
        const int S2 = (int)std::max<long long>(1, std::min<long long>(Smax, std::max(S, Sw)));
        ksplit = (Kc + S2 - 1) / S2;
        S = (Kc + ksplit - 1) / ksplit;
    }
    dim3 grid((ne + 255) / 256, S, groups);
    {
        StageScope sc(c, ST_CONTRACT, Kc);
        const size_t nph = (size_t)Kc * ldp;
        hipLaunchKernelGGL(lqm::k_phase_table, dim3((unsigned)((nph + 255) / 256)), dim3(256), 0, s, Kc, d, ks, kw,
                           wuni, nRr, ldp, Rr, -1.0, ph);
#define LQM_CONTRACT(G)                                                                                        \
    hipLaunchKernelGGL(lqm::k_phase_contract_s<G>, grid, dim3(256), 0, s, NN, ne, Nh, Kc, ph, ldp, rhok, ksplit, \
                       nRr, part)
        switch (RG) {
            case 1: LQM_CONTRACT(1); break;
            case 4: LQM_CONTRACT(4); break;
            case 8: LQM_CONTRACT(8); break;
            case 9: LQM_CONTRACT(9); break;
            case 16: LQM_CONTRACT(16); break;
            case 24: LQM_CONTRACT(24); break;
            case 25: LQM_CONTRACT(25); break;
            case 28: LQM_CONTRACT(28); break;
            default: LQM_CONTRACT(32); break;
        }
#undef LQM_CONTRACT
    }
    const size_t total = (size_t)nRr * (herm ? (size_t)ne : (size_t)NN);
    StageScope sc(c, ST_FOLD);
    const dim3 gf((unsigned)((total + 255) / 256));
#define LQM_FOLD(G)                                                                                       \
    do {                                                                                                  \
        if (herm)                                                                                         \
            hipLaunchKernelGGL(lqm::k_phase_fold_h<G>, gf, dim3(256), 0, s, N, nRr, S, neg, part, rho); \
        else                                                                                              \
            hipLaunchKernelGGL(lqm::k_phase_fold<G>, gf, dim3(256), 0, s, NN, nRr, S, part, rho);       \
    } while (0)
    switch (RG) {
        case 1: LQM_FOLD(1); break;
        case 4: LQM_FOLD(4); break;
        case 8: LQM_FOLD(8); break;
        case 9: LQM_FOLD(9); break;
        case 16: LQM_FOLD(16); break;
        case 24: LQM_FOLD(24); break;
        case 25: LQM_FOLD(25); break;
        case 28: LQM_FOLD(28); break;
        default: LQM_FOLD(32); break;
    }
#undef LQM_FOLD
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

lqm_status check_status(lqm_context* c, const char* what) {
    int h = 0;
    LQM_HIP(hipMemcpyAsync(&h, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LQM_HIP(hipStreamSynchronize(c->stream));
    if (h) {
        LQM_HIP(hipMemsetAsync(c->d_status, 0, sizeof(int), c->stream));
        if (h & lqm::kStatNoConv)
            return fail(LQM_ERR_NOCONV, std::string(what) + ": tridiagonal QL did not converge");
        if (h & lqm::kStatLogFull)
            return fail(LQM_ERR_NOCONV, std::string(what) + ": QL rotation log capacity exceeded");
        if (h & lqm::kStatNonFinite)
            return fail(LQM_ERR_NONFINITE, std::string(what) + ": non-finite eigenvalue (NaN/Inf in H(k))");
    }
    return LQM_OK;
}

lqm_status check_model(lqm_model* m) {
    if (!m || !m->ctx) return fail(LQM_ERR_ARG, "null model");
    return LQM_OK;
}

// bands (+ evecs) over the whole grid, device pointers.
lqm_status run_mixed(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                     int nRr, const int* dRr, cx<double>* rho, double* eband, double* bands,
                     const double* occ = nullptr);

// Chemical potential requested from the pass itself: when the whole grid is one chunk, mu
// is solved (filling, the pass's own eigenvalues) between the eigenvalues and the
// occupations, and written to *mu; otherwise the pass returns kFallback untouched.  With
// dres (T > 0, per-k kernels) mu stays on the device: the bisection writes dres[0] = mu,
// dres[1] = its status, the occupations are formed from dres[0], and the pass neither
// writes *mu nor reads the status word -- the caller reads dres and the status with its
// own next device-to-host copy.
struct MuReq {
    double filling;
    double* mu;
    double* dres = nullptr;
    mutable bool on_device = false;  // set by the pass when it took the dres route
};
// internal status, never returned through the C ABI: take the two-pass route (7 lies in the
// enumeration's value range, above every public code)
constexpr lqm_status kFallback = static_cast<lqm_status>(7);
lqm_status mu_from_bands(lqm_context* c, int N, int Nk, const double* bands, const double* kw,
                         double filling, double T, double* mu);
lqm_status run_blocked(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                       int nRr, const int* dRr, cx<double>* rho, double* eband, double* bands,
                       double* evecs, const double* occ, const MuReq* mq = nullptr);
bool use_blocked(const lqm_context* c, int N);
lqm_status run_blocked_mixed(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                             int nRr, const int* dRr, cx<double>* rho, double* eband, double* bands,
                             const double* occ, const MuReq* mq);

// Start of a pass over the grid: clears the refinement report of the previous pass and
// records the precision this pass runs in (reported by lqm_context_last_precision).
void begin_pass(lqm_context* c, int prec) {
    for (double& r : c->report) r = 0.0;
    c->last_prec = prec;
}

// Clears the context's occupation pointer when a pass ends, on every return path.
struct OccScope {
    lqm_context* c;
    ~OccScope() { c->occ = nullptr; }
};

// Models of at most four orbitals: one thread per k-point (k_small_eig / k_small_rho) in
// place of the per-k workgroup kernels, unless a variant of those kernels is forced.
bool use_small(const lqm_context* c, int N) {
    if (N > 4 || c->precision != LQM_PREC_FP64) return false;
    for (int k = 0; k < VK_N; ++k)
        if (c->forced[k] >= 0) return false;
    return true;
}
lqm_status launch_small_eig(lqm_model* m, int Kc, const double* ks, double* ev, cx<double>* V) {
    lqm_context* c = m->ctx;
    StageScope sc(c, ST_TRIDIAG, Kc);
    const dim3 g((Kc + 255) / 256);
    switch (m->N) {
        case 1: hipLaunchKernelGGL(lqm::k_small_eig<1>, g, dim3(256), 0, c->stream, Kc, m->d, m->nR, m->dR, m->dT, ks, ev, V, c->d_status); break;
        case 2: hipLaunchKernelGGL(lqm::k_small_eig<2>, g, dim3(256), 0, c->stream, Kc, m->d, m->nR, m->dR, m->dT, ks, ev, V, c->d_status); break;
        case 3: hipLaunchKernelGGL(lqm::k_small_eig<3>, g, dim3(256), 0, c->stream, Kc, m->d, m->nR, m->dR, m->dT, ks, ev, V, c->d_status); break;
        default: hipLaunchKernelGGL(lqm::k_small_eig<4>, g, dim3(256), 0, c->stream, Kc, m->d, m->nR, m->dR, m->dT, ks, ev, V, c->d_status); break;
    }
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}
lqm_status launch_small_rho(lqm_context* c, int N, int Kc, const double* ev, const cx<double>* V, const double* kw,
                            double wuni, double mu, const double* mu_dev, double T, const double* occ,
                            cx<double>* rhok, double* eb) {
    StageScope sc(c, ST_BACKTRANSFORM, Kc);
    const dim3 g((Kc + 255) / 256);
    switch (N) {
        case 1: hipLaunchKernelGGL(lqm::k_small_rho<1>, g, dim3(256), 0, c->stream, Kc, ev, V, kw, wuni, mu, mu_dev, T, occ, rhok, eb); break;
        case 2: hipLaunchKernelGGL(lqm::k_small_rho<2>, g, dim3(256), 0, c->stream, Kc, ev, V, kw, wuni, mu, mu_dev, T, occ, rhok, eb); break;
        case 3: hipLaunchKernelGGL(lqm::k_small_rho<3>, g, dim3(256), 0, c->stream, Kc, ev, V, kw, wuni, mu, mu_dev, T, occ, rhok, eb); break;
        default: hipLaunchKernelGGL(lqm::k_small_rho<4>, g, dim3(256), 0, c->stream, Kc, ev, V, kw, wuni, mu, mu_dev, T, occ, rhok, eb); break;
    }
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

lqm_status run_bands(lqm_model* m, int Nk, const double* ks, double* bands, double* evecs) {
    const int N = m->N;
    if (Nk < 0) return fail(LQM_ERR_ARG, "Nk < 0");
    if (Nk == 0) return LQM_OK;
    if (!ks || !bands) return fail(LQM_ERR_ARG, "null ks or bands");
    if (use_blocked(m->ctx, N)) {
        if (m->ctx->precision == LQM_PREC_FP32_REFINED && !evecs)
            return run_blocked_mixed(m, Nk, ks, nullptr, 0.0, 0.0, 0, nullptr, nullptr, nullptr, bands, nullptr, nullptr);
        return run_blocked(m, Nk, ks, nullptr, 0.0, 0.0, 0, nullptr, nullptr, nullptr, bands, evecs, nullptr);
    }
    // eigenvector output always comes from the FP64 path
    if (m->ctx->precision == LQM_PREC_FP32_REFINED && !evecs)
        return run_mixed(m, Nk, ks, nullptr, 0.0, 0.0, 0, nullptr, nullptr, nullptr, bands);
    begin_pass(m->ctx, LQM_PREC_FP64);
    if (use_small(m->ctx, N)) {
        lqm_status st = launch_small_eig(m, Nk, ks, bands, reinterpret_cast<cx<double>*>(evecs));
        if (st) return st;
        return check_status(m->ctx, "bandmatrix");
    }
    Plan P = make_plan(m->ctx, N, m->nR, Nk, evecs != nullptr, false);
    for (int k0 = 0; k0 < Nk; k0 += P.Kc) {
        const int Kc = std::min(P.Kc, Nk - k0);
        const double* ksc = ks + (size_t)k0 * m->d;
        double* ev = bands + (size_t)k0 * N;
        lqm_status st = launch_eigvals(m, P, Kc, ksc, ev, evecs != nullptr);
        if (st) return st;
        if (evecs) {
            st = launch_backtransform<double, 0>(m, P, Kc, ev, nullptr, 0.0, 0.0, 0.0,
                                         reinterpret_cast<cx<double>*>(evecs) + (size_t)k0 * N * N,
                                         nullptr);
            if (st) return st;
        }
    }
    return check_status(m->ctx, "bandmatrix");
}

// occ (device, N x Nk, may be null): occupations of the ascending eigenvalues used in
// place of fermidirac(e - mu; T) -- rho then is sum_k w_k e^{-2 pi i k.R} U diag(occ_k) U^H.
// dRr_dev (optional): the keys already on the device (no upload: the SCF's captured passes).
// Whether the key list (host, d components each) has no repeated key and holds the negation
// of every key: then R -> -R is a bijection of the list, which the Hermitian contraction needs.
bool keys_closed_under_negation(int nRr, int d, const int32_t* Rr) {
    if (nRr <= 0 || !Rr) return false;
    for (int r = 0; r < nRr; ++r) {
        int found = 0;
        for (int q = 0; q < nRr; ++q) {
            bool neg = true, same = q != r;
            for (int a = 0; a < d; ++a) {
                neg = neg && Rr[q * d + a] == -Rr[r * d + a];
                same = same && Rr[q * d + a] == Rr[r * d + a];
            }
            if (same) return false;
            found += neg;
        }
        if (found != 1) return false;
    }
    return true;
}
struct PairsScope {
    lqm_context* c;
    ~PairsScope() { c->rr_pairs = 0; }
};

lqm_status run_rho(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                   int nRr, const int32_t* Rr, double* rho, double* eband, double* bands_out,
                   const double* occ = nullptr, const MuReq* mq = nullptr, const int* dRr_dev = nullptr) {
    lqm_context* c = m->ctx;
    if (mq && !use_blocked(c, m->N) && c->precision == LQM_PREC_FP32_REFINED)
        return kFallback;  // the refined eigenvalues come with the occupations in one kernel
    const int N = m->N, d = m->d;
    if (Nk <= 0) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (nRr < 0 || (nRr > 0 && !Rr)) return fail(LQM_ERR_ARG, "bad rho key list");
    if (!ks || !rho || !eband) return fail(LQM_ERR_ARG, "null ks, rho or eband");
    const size_t NN = (size_t)N * N;
    c->rr_pairs = keys_closed_under_negation(nRr, d, Rr);
    PairsScope pairs_scope{c};
    zero_dev(c, rho, sizeof(cx<double>) * NN * std::max(nRr, 0));
    zero_dev(c, eband, sizeof(double));
    const int* dRr = dRr_dev;
    lqm_status st;
    if (!dRr) {
        int* up = nullptr;
        if ((st = wsT(c, W_RR, (size_t)std::max(1, nRr * d), &up))) return st;
        if (nRr > 0)
            LQM_HIP(hipMemcpyAsync(up, Rr, sizeof(int) * nRr * d, hipMemcpyHostToDevice, c->stream));
        dRr = up;
    }
    if (use_blocked(c, N)) {
        if (c->precision == LQM_PREC_FP32_REFINED)
            return run_blocked_mixed(m, Nk, ks, kw, mu, T, nRr, dRr, reinterpret_cast<cx<double>*>(rho), eband,
                                     bands_out, occ, mq);
        return run_blocked(m, Nk, ks, kw, mu, T, nRr, dRr, reinterpret_cast<cx<double>*>(rho), eband,
                           bands_out, nullptr, occ, mq);
    }
    if (c->precision == LQM_PREC_FP32_REFINED)
        return run_mixed(m, Nk, ks, kw, mu, T, nRr, dRr, reinterpret_cast<cx<double>*>(rho), eband,
                         bands_out, occ);
    Plan P = make_plan(c, N, m->nR, Nk, true, true);
    if (mq && P.Kc < Nk) return kFallback;  // mu needs every eigenvalue before the first occupation
    begin_pass(c, LQM_PREC_FP64);
    const int RG = contract_group(nRr);
    const int groups = std::max(1, (nRr + RG - 1) / RG);
    const int tilesE = (int)((NN + 255) / 256);
    int S = std::max(1, std::min(std::max(1, P.Kc / kContractMinK), contract_splits(tilesE * groups)));
    cx<double>* part = nullptr;
    if ((st = wsT(c, W_PART, (size_t)groups * S * RG * NN, &part))) return st;
    const double wuni = 1.0 / (double)Nk;
    OccScope occ_scope{c};
    for (int k0 = 0; k0 < Nk; k0 += P.Kc) {
        c->occ = occ ? occ + (size_t)k0 * N : nullptr;
        const int Kc = std::min(P.Kc, Nk - k0);
        const double* ksc = ks + (size_t)k0 * d;
        const double* kwc = kw ? kw + k0 : nullptr;
        double* ev = nullptr;
        if (bands_out) {
            ev = bands_out + (size_t)k0 * N;
        } else if ((st = wsT(c, W_EVALS, (size_t)N * P.Kc, &ev))) {
            return st;
        }
        const bool small = use_small(c, N);
        cx<double>* Vs = nullptr;
        if (small) {
            if ((st = wsT(c, W_SMV, NN * P.Kc, &Vs))) return st;
            if ((st = launch_small_eig(m, Kc, ksc, ev, Vs))) return st;
        } else if ((st = launch_eigvals(m, P, Kc, ksc, ev, true))) {
            return st;
        }
        if (mq && mq->dres && T > 0.0) {
            double *focc, *ework = nullptr;
            if ((st = wsT(c, W_FOCC, (size_t)N * Kc, &focc))) return st;
            if ((size_t)N * Nk > 16 * 1024 && (st = wsT(c, W_MUE, (size_t)N * Nk, &ework))) return st;
            {
                StageScope sc(c, ST_MU);
                hipLaunchKernelGGL(lqm::k_mu_bisect, dim3(1), dim3(1024), 0, c->stream, N, N * Nk, ev, kw,
                                   mq->filling, T, mq->dres, ework);
                // the one-thread-per-k-point path forms fermidirac(e - mu) itself from the device mu
                const int L = N * Kc;
                if (!small)
                    hipLaunchKernelGGL(lqm::k_fermi_occ, dim3((L + 255) / 256), dim3(256), 0, c->stream, L, ev,
                                       mq->dres, T, focc);
            }
            LQM_HIP(hipGetLastError());
            if (!small) c->occ = focc;
            mq->on_device = true;
        } else if (mq) {
            if ((st = mu_from_bands(c, N, Nk, ev, kw, mq->filling, T, &mu))) return st;
            *mq->mu = mu;
        }
        cx<double>* rhok;
        double* eb;
        if ((st = wsT(c, W_RHOK, NN * P.Kc, &rhok))) return st;
        if ((st = wsT(c, W_EB, (size_t)P.Kc, &eb))) return st;
        if (small) {
            const double* mud = mq && mq->on_device ? mq->dres : nullptr;
            if ((st = launch_small_rho(c, N, Kc, ev, Vs, kwc, wuni, mu, mud, T, c->occ, rhok, eb))) return st;
        } else if (P.rho2) {
            if ((st = launch_rho2(m, P, Kc, ev, kwc, wuni, mu, T, rhok, eb))) return st;
        } else if ((st = launch_backtransform<double, 1>(m, P, Kc, ev, kwc, wuni, mu, T, rhok, eb))) {
            return st;
        }
        {
            StageScope sc(c, ST_REDUCE);
            hipLaunchKernelGGL(lqm::k_sum, dim3(1), dim3(256), 0, c->stream, Kc, eb, eband, 1);
        }
        if (nRr > 0) {
            const int Sc = std::min(S, Kc);
            const int ksplit = (Kc + Sc - 1) / Sc;
            const int Sused = (Kc + ksplit - 1) / ksplit;
            auto* rh = reinterpret_cast<cx<double>*>(rho);
            if ((st = launch_contract(RG, c, N, Kc, d, ksc, kwc, wuni, nRr, dRr, rhok, ksplit, Sused, groups, part, rh))) return st;
            LQM_HIP(hipGetLastError());
        }
    }
    if (mq && mq->on_device) return LQM_OK;  // the caller reads the status word with dres
    return check_status(c, "densitymatrix");
}

// ---- mixed precision: FP32 eigensolve + FP64 Ogita-Aishima refinement ---------------
// C = alpha op(A) op(B) + beta D over nb k-points (kidx: their indices, null = 0..nb-1).
template <int OPA, int OPB>
void zgemm(lqm_context* c, int N, int nb, const cx<double>* A, const cx<double>* B, cx<double>* C,
           cx<double> alpha, cx<double> beta, const cx<double>* D = nullptr,
           const int* kidx = nullptr) {
    StageScope sc(c, OPA == 0 ? (OPB == 0 ? ST_GEMM_NN : ST_GEMM_NC) : ST_GEMM_CN, nb);
    const int t = (N + lqm::kGT - 1) / lqm::kGT;
    // op(A) = A: matrix-core kernel (fastest measured); products with a conjugate
    // transpose: the LDS-tiled VALU kernel, whose staging reads them coalesced
    if (OPA == 0 && OPB == 0) {
        auto kG = lqm::k_zgemm_mfma<OPA, OPB>;
        hipLaunchKernelGGL(kG, dim3(t, t, nb), dim3(256), 0, c->stream, N, A, B, C, D, alpha, beta, kidx);
    } else {
        auto kG = lqm::k_zgemm_batched<OPA, OPB>;
        hipLaunchKernelGGL(kG, dim3(t, t, nb), dim3(256), 0, c->stream, N, A, B, C, D, alpha, beta, kidx);
    }
}

// One pass over the grid.  Outputs (any may be null): sorted refined eigenvalues
// (bands), rho on keys dRr (rho: accumulated into, zeroed by the caller) and eband.
lqm_status run_mixed(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                     int nRr, const int* dRr, cx<double>* rho, double* eband, double* bands,
                     const double* occ) {
    lqm_context* c = m->ctx;
    const int N = m->N, d = m->d;
    if (N > 512) return fail(LQM_ERR_UNSUPPORTED, "mixed precision path supports N <= 512");
    begin_pass(c, LQM_PREC_FP32_REFINED);
    Plan P = make_plan(c, N, m->nR, Nk, true, false, sizeof(float));
    const size_t NN = (size_t)N * N;
    // refinement buffers: X, A, W, S, G (+ rho_k) in FP64 complex per k-point, X in FP32
    const size_t per64 = 16 * NN * (rho ? 6 : 5) + 8 * NN + 48 * N;
    P.Kc = (int)std::max<size_t>(1, std::min<size_t>((size_t)P.Kc, std::min(kWorkspaceBudget, c->ws_cap) / per64));
    const double wuni = 1.0 / (double)Nk;
    const cx<double> one = {1.0, 0.0}, zero = {0.0, 0.0};
    lqm_status st;
    int sweeps_max = 0;
    double rmax = 0.0, dmax = 0.0, nnear = 0.0;
    const int RG = contract_group(nRr);
    const int groups = std::max(1, (nRr + RG - 1) / RG);
    const int tilesE = (int)((NN + 255) / 256);
    const int S = std::max(1, std::min(std::max(1, P.Kc / kContractMinK), contract_splits(tilesE * groups)));
    cx<double>* part = nullptr;
    if (rho && nRr > 0 && (st = wsT(c, W_PART, (size_t)groups * S * RG * NN, &part))) return st;
    for (int k0 = 0; k0 < Nk; k0 += P.Kc) {
        const int Kc = std::min(P.Kc, Nk - k0);
        const double* ksc = ks + (size_t)k0 * d;
        const double* kwc = kw ? kw + k0 : nullptr;
        float* ev32;
        cx<double> *X, *A, *W, *Sm, *G;
        double *lam, *stat;
        if ((st = wsT(c, W_EV32, (size_t)N * P.Kc, &ev32))) return st;
        if ((st = wsT(c, W_X, NN * P.Kc, &X))) return st;
        if ((st = wsT(c, W_HA, NN * P.Kc, &A))) return st;
        if ((st = wsT(c, W_W, NN * P.Kc, &W))) return st;
        if ((st = wsT(c, W_S, NN * P.Kc, &Sm))) return st;
        if ((st = wsT(c, W_G, NN * P.Kc, &G))) return st;
        if ((st = wsT(c, W_LAM, (size_t)N * P.Kc, &lam))) return st;
        if ((st = wsT(c, W_STAT, (size_t)4 * P.Kc, &stat))) return st;
        // H(k) in FP64 (the refinement's A), then FP32 eigenvectors, ascending, as FP64 complex
        {
            StageScope sc(c, ST_HK64, Kc);
            if ((st = launch_hk_full(c, m, Kc, ksc, A))) return st;
        }
        LQM_HIP(hipGetLastError());
        if ((st = launch_eigvals<float>(m, P, Kc, ksc, ev32, true, A))) return st;
        // refinement kind: 2 = one fused kernel per k-point and sweep in column-panel form
        // (two blocks per CU; the default for 16 < N <= 64, as it works on a 64 x 64 tile),
        // 1 = the same with X and W in LDS (one block per CU), 0 = the GEMM chain below
        const int fk = c->forced[VK_REFINE] >= 0 ? c->forced[VK_REFINE] : (N > 16 ? 2 : 0);
        const bool fused = N <= 64 && fk >= 1;
        // the panel sweep widens FP32 eigenvectors itself: the compact WY then writes them in
        // FP32 (half the bytes each way), and the first sweep writes FP64 X where needed
        cx<float>* X32 = nullptr;
        bool x32 = false;
        if (fused && fk == 2 && (st = wsT(c, W_X32, NN * P.Kc, &X32))) return st;
        if ((st = launch_backtransform<float, 0>(m, P, Kc, ev32, nullptr, 0.0, 0.0, 0.0, X, nullptr, X32, &x32)))
            return st;
        // Ogita-Aishima sweeps.  The first runs on every k-point; later ones only on the
        // k-points whose last correction flagged another sweep (a rotated or unresolved
        // cluster, or a correction whose square -- the error it leaves -- is above the
        // threshold), addressed through a device index list.
        int sweeps = 0;
        int nsub = Kc;
        const int* kidx = nullptr;
        const int max_sw = std::max(2, c->refine_max);
        // the bookkeeping of each sweep stays on the device (k_oa_scan): the k-point lists
        // [cluster | next sweep (two buffers, read and written alternately) | rho_k not formed]
        // and a header, of which only the header comes back
        int* scan_l = nullptr;
        lqm::OaScanHdr* dhdr = nullptr;
        if ((st = wsT(c, W_OASCAN, (size_t)4 * Kc, &scan_l))) return st;
        if ((st = wsT(c, W_OAHDR, (size_t)1, &dhdr))) return st;
        if (!c->pin) LQM_HIP(hipHostMalloc(&c->pin, sizeof(lqm::OaScanHdr), hipHostMallocDefault));
        int lbuf = 0;  // the next-sweep buffer the scan writes
        lqm::OaScanHdr h{};
        auto scan = [&](const int* curl, int ncur) -> lqm_status {
            hipLaunchKernelGGL(lqm::k_oa_scan, dim3(1), dim3(lqm::kScanT), 0, c->stream, Kc, stat, curl, ncur, dhdr,
                               scan_l, scan_l + (size_t)(1 + lbuf) * Kc, scan_l + (size_t)3 * Kc);
            LQM_HIP(hipGetLastError());
            LQM_HIP(hipMemcpyAsync(c->pin, dhdr, sizeof(lqm::OaScanHdr), hipMemcpyDeviceToHost, c->stream));
            LQM_HIP(hipStreamSynchronize(c->stream));
            h = *static_cast<const lqm::OaScanHdr*>(c->pin);
            return LQM_OK;
        };
        // with rho or eband requested, a fused sweep forms rho_k (into Sm) and eb of every
        // k-point it finishes; the rest go through k_scale_occ and the NC product below
        double* eb = nullptr;
        if ((rho || eband) && (st = wsT(c, W_EB, (size_t)P.Kc, &eb))) return st;
        const double* occk = occ ? occ + (size_t)k0 * N : nullptr;
        // a key list closed under R -> -R: the contraction reads rho_k on and above the diagonal
        const lqm::OaRho rq{(rho || eband) ? 1 : 0, kwc, wuni, mu, T, occk, Sm, eb,
                            (rho && c->rr_pairs && nRr > 0) ? 1 : 0};
        for (;;) {
            if (fused) {
                {
                // the first sweep (every k-point) and the later ones (listed k-points) timed apart
                StageScope sc(c, kidx ? ST_OA2 : ST_OA, nsub);
                if (fk == 2 && x32) {
                    // first sweep, every k-point: X read in FP32 from the compact WY
                    const size_t sh = lqm::oa_panel_lds_bytes();
                    allow_lds(lqm::k_oa_panel<cx<float>>, sh);
                    hipLaunchKernelGGL(lqm::k_oa_panel<cx<float>>, dim3(nsub), dim3(256), sh, c->stream, N, A,
                                       static_cast<const cx<float>*>(X32), X, Sm, G, lam, stat, c->refine_thr,
                                       kidx, rq);
                    x32 = false;
                } else if (fk == 2) {
                    const size_t sh = lqm::oa_panel_lds_bytes();
                    allow_lds(lqm::k_oa_panel<cx<double>>, sh);
                    hipLaunchKernelGGL(lqm::k_oa_panel<cx<double>>, dim3(nsub), dim3(256), sh, c->stream, N, A,
                                       static_cast<const cx<double>*>(X), X, Sm, G, lam, stat, c->refine_thr, kidx, rq);
                } else {
                    const size_t sh = lqm::oa_fused_lds_bytes();
                    allow_lds(lqm::k_oa_fused, sh);
                    hipLaunchKernelGGL(lqm::k_oa_fused, dim3(nsub), dim3(256), sh, c->stream, N, A, X, Sm, G, lam,
                                       stat, c->refine_thr, kidx, rq);
                }
                }
                LQM_HIP(hipGetLastError());
                if ((st = scan(kidx, nsub))) return st;
                // k-points with a cluster: general correction (Rayleigh-Ritz) on their S, G
                if (h.n8 > 0) {
                    const int* dk2 = scan_l;
                    const int nc = h.n8;
                    {
                        StageScope sc(c, ST_OA2, nc);
                        const size_t sh = lqm::oa_lds_bytes(N);
                        allow_lds(lqm::k_oa_correct, sh);
                        hipLaunchKernelGGL(lqm::k_oa_correct, dim3(nc), dim3(lqm::kOaT), sh, c->stream, N, Sm, G, W,
                                           lam, stat, c->refine_thr, dk2);
                    }
                    LQM_HIP(hipGetLastError());
                    zgemm<0, 0>(c, N, nc, X, W, G, one, one, X, dk2);
                    const int bx = (int)std::min<size_t>(64, (NN + 255) / 256);
                    hipLaunchKernelGGL(lqm::k_copy_sub, dim3(bx, nc), dim3(256), 0, c->stream, (int)NN, nc, dk2, G, X);
                    LQM_HIP(hipGetLastError());
                    // the correction rewrote those k-points' status rows
                    if ((st = scan(kidx, nsub))) return st;
                }
            } else {
            zgemm<0, 0>(c, N, nsub, A, X, W, one, zero, nullptr, kidx);   // W = A X
            zgemm<1, 0>(c, N, nsub, X, W, Sm, one, zero, nullptr, kidx);  // S = X^H A X
            zgemm<1, 0>(c, N, nsub, X, X, G, one, zero, nullptr, kidx);   // G = X^H X
            {
                // the first sweep (every k-point) and the later ones (listed k-points) timed apart
                StageScope sc(c, kidx ? ST_OA2 : ST_OA, nsub);
                const size_t sh = lqm::oa_lds_bytes(N);
                allow_lds(lqm::k_oa_correct, sh);
                hipLaunchKernelGGL(lqm::k_oa_correct, dim3(nsub), dim3(lqm::kOaT), sh, c->stream, N, Sm, G, W, lam,
                                   stat, c->refine_thr, kidx);
            }
            LQM_HIP(hipGetLastError());
            // X <- X + X E  (E in W), written to G; then G holds the refined vectors
            zgemm<0, 0>(c, N, nsub, X, W, G, one, one, X, kidx);
            if (kidx) {
                const int bx = (int)std::min<size_t>(64, (NN + 255) / 256);
                hipLaunchKernelGGL(lqm::k_copy_sub, dim3(bx, nsub), dim3(256), 0, c->stream, (int)NN, nsub, kidx,
                                   G, X);
                LQM_HIP(hipGetLastError());
            } else {
                std::swap(X, G);
            }
            if ((st = scan(kidx, nsub))) return st;
            }
            ++sweeps;
            rmax = std::max(rmax, h.rmax);
            dmax = std::max(dmax, h.dmax);
            if (h.nonfinite) return fail(LQM_ERR_NONFINITE, "mixed precision: non-finite refinement residual");
            c->report[4] += (double)nsub;
            if (h.n4 == 0) break;
            if (sweeps >= max_sw) {
                nnear += h.unresolved;
                if (h.unresolved)
                    return fail(LQM_ERR_NOCONV, "mixed precision: a near-degenerate cluster wider than 8 "
                                                "eigenvalues could not be resolved; use the FP64 path");
                break;
            }
            kidx = scan_l + (size_t)(1 + lbuf) * Kc;
            nsub = h.n4;
            lbuf ^= 1;
        }
        sweeps_max = std::max(sweeps_max, sweeps);
        if (bands) {
            StageScope sc(c, ST_REDUCE);
            // one wave per k-point up to 64 eigenvalues (the kernel strides by its block size)
            hipLaunchKernelGGL(lqm::k_sort_evals, dim3(Kc), dim3(std::min(256, (N + 63) / 64 * 64)), 0, c->stream,
                               N, lam, bands + (size_t)k0 * N);
        }
        if (rho || eband) {
            // k-points whose rho_k the fused sweep did not form (all of them on the GEMM
            // chain; clusters and k-points stopped at the sweep limit on the fused path)
            int nrest = Kc;
            const int* krest = nullptr;
            if (fused) {
                nrest = h.nrest;  // from the last sweep's scan, over every k-point of the chunk
                c->report[5] += (double)(Kc - nrest);
                if (nrest > 0 && nrest < Kc) krest = scan_l + (size_t)3 * Kc;
            }
            {
                StageScope sc(c, ST_REDUCE);
                if (nrest > 0)
                    hipLaunchKernelGGL(lqm::k_scale_occ, dim3(nrest), dim3(256), sizeof(double) * N, c->stream,
                                       N, X, lam, kwc, wuni, mu, T, occk, G, eb, krest);
                hipLaunchKernelGGL(lqm::k_sum, dim3(1), dim3(256), 0, c->stream, Kc, eb, eband, 1);
            }
            if (nrest > 0) zgemm<0, 1>(c, N, nrest, G, X, Sm, one, zero, nullptr, krest);  // rho_k = X f X^H
            if (rho && nRr > 0) {
                const int Sc = std::min(S, Kc);
                const int ksplit = (Kc + Sc - 1) / Sc;
                const int Sused = (Kc + ksplit - 1) / ksplit;
                if ((st = launch_contract(RG, c, N, Kc, d, ksc, kwc, wuni, nRr, dRr, Sm, ksplit, Sused, groups, part, rho))) return st;
            }
            LQM_HIP(hipGetLastError());
        }
    }
    c->report[0] = sweeps_max;
    c->report[1] = rmax;
    c->report[2] = dmax;
    c->report[3] = nnear;
    return check_status(c, "mixed-precision pass");
}

// ---- large N: blocked tridiagonalization, bisection, inverse iteration, compact WY ------
constexpr size_t kBlockedBudget = size_t(40) << 30;  // bytes of per-k scratch per chunk
// default: the blocked path above these N (measured crossovers of the per-k kernels and the
// large-N path with the column-per-lane reduction, 4096 k-points of a 9-key model)
constexpr int kBlockedMinN = 64;         // FP64
constexpr int kBlockedMinNMixed = 64;    // FP32 with refinement
constexpr double kClusterRel = 1e-5;  // eigenvalues closer than this x ||T|| share a cluster
constexpr double kOccTiny = 1e-18;    // occupations below this leave rho_k out

bool use_blocked(const lqm_context* c, int N) {
    const int f = c->forced[VK_BLOCKED];
    if (f >= 0) return f == 1;
    auto it = c->tuned.find((VK_BLOCKED << 20) | N);
    if (it != c->tuned.end()) return it->second == 1;
    return N > (c->precision == LQM_PREC_FP32_REFINED ? kBlockedMinNMixed : kBlockedMinN);
}

template <int OPA, int OPB, typename R = double>
void zgemm_gen(lqm_context* c, int stage, long long kpts, int nb, const lqm::GemmArgs<R>& g) {
    if (g.M <= 0 || g.N <= 0 || nb <= 0) return;
    StageScope sc(c, stage, kpts);
    const unsigned nt = (unsigned)((g.M + 63) / 64);
    const dim3 grid = g.lower ? dim3(nt * (nt + 1) / 2, 1, (unsigned)nb)
                              : dim3(nt, (unsigned)((g.N + 63) / 64), (unsigned)nb);
    hipLaunchKernelGGL((lqm::k_zgemm_gen<OPA, OPB, R>), grid, dim3(256), 0, c->stream, g);
}

template <typename R = double>
lqm::GemmArgs<R> gemm_args(int M, int N, int K, const cx<R>* A, int lda, long long sA,
                           const cx<R>* B, int ldb, long long sB, cx<R>* C, int ldc, long long sC,
                           cx<double> alpha, cx<double> beta) {
    lqm::GemmArgs<R> g{};
    g.M = M;
    g.N = N;
    g.K = K;
    g.npair = 1;
    g.A[0] = A;
    g.B[0] = B;
    g.A[1] = nullptr;
    g.B[1] = nullptr;
    g.lda = lda;
    g.ldb = ldb;
    g.ldc = ldc;
    g.sA = sA;
    g.sB = sB;
    g.sC = sC;
    g.sS = 0;
    g.C = C;
    g.alpha = alpha;
    g.beta = beta;
    g.ascale = nullptr;
    g.kidx = nullptr;
    g.pairs = nullptr;
    return g;
}

// zhetrd (lower) of Kc full Hermitian matrices A with zlatrd panels: d, e, tau out, the
// reflectors left in A below the diagonal.  R: working precision (the tridiagonal d, e is
// written in FP64 either way).
constexpr int kTrdSplitK = 32;  // k-points per half of a split tridiagonalization, at least
template <typename R>
lqm_status blocked_tridiag(lqm_context* c, int N, int Kc, cx<R>* A, cx<R>* Vp, cx<R>* Wp,
                           cx<R>* P, cx<R>* P2, cx<R>* Q, double* SS, int nchmax, int nbmax,
                           double* dd, double* ee, cx<R>* tau) {
    hipStream_t s = c->stream;
    const size_t NN = (size_t)N * N;
    if (N > 64 && N <= 128 && c->forced[VK_TRIDIAG] != 0) {
        // the column-per-lane reduction in registers (512 threads per k-point), the same outputs;
        // variant 0 of kind 0 keeps the zlatrd panels below
        StageScope sc(c, ST_BTRD, Kc);
        hipLaunchKernelGGL(lqm::k_tridiag_cg<R>, dim3(Kc), dim3(512), 0, s, N, A, dd, ee, tau);
        LQM_HIP(hipGetLastError());
        return LQM_OK;
    }
    // Otherwise zlatrd panels.  A chunk of at least 2 kTrdSplitK k-points is cut in two
    // halves whose column steps go to the context stream and the side stream, interleaved
    // column by column, so that one half's latency-bound steps (k_trd_col, k_trd_w) run beside
    // the other half's bandwidth-bound k_trd_hemv (every buffer is per k-point, so a half is
    // the same call on offset pointers).  Not with stage timing on, which times each stage on
    // one stream.
    const bool split = !c->timing && c->side && Kc >= 2 * kTrdSplitK;
    const int nh = split ? 2 : 1;
    const int Kh[2] = {split ? Kc / 2 : Kc, split ? Kc - Kc / 2 : 0};
    const size_t k0[2] = {0, (size_t)Kh[0]};
    hipStream_t st_of[2] = {c->stream, c->side};
    const hipStream_t s_main = c->stream;
    if (split) {
        if (c->tfac_pending) {
            LQM_HIP(hipStreamWaitEvent(s_main, c->ev_join, 0));
            c->tfac_pending = false;
        }
        LQM_HIP(hipEventRecord(c->ev_fork, s_main));
        LQM_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    }
    const size_t NP = (size_t)N * lqm::kBnb;
    auto col_step = [&](int h, int j, int i, int last) {
        const size_t o = k0[h];
        hipLaunchKernelGGL(lqm::k_trd_col<R>, dim3(Kh[h]), dim3(lqm::kBColT), 0, st_of[h], N, j, i, last, A + o * NN,
                           Vp + o * NP, Wp + o * NP, dd + o * N, ee + o * N, tau + o * N);
        if (last) return;
        const int n = N - j - 1;
        const int nch = (n + lqm::kHvCols - 1) / lqm::kHvCols;
        hipLaunchKernelGGL(lqm::k_trd_hemv<R>, dim3(lqm::hv_blocks(nch), Kh[h]), dim3(lqm::kHvWaves * 64), 0, st_of[h],
                           N, j, i, nchmax, nbmax, A + o * NN, Vp + o * NP, Wp + o * NP, P + o * nchmax * N,
                           P2 + o * nchmax * N, Q + o * nchmax * lqm::kQStride, SS + o * nbmax);
        hipLaunchKernelGGL(lqm::k_trd_w<R>, dim3((n + 255) / 256, Kh[h]), dim3(256), 0, st_of[h], N, j, i, nchmax,
                           nbmax, P + o * nchmax * N, P2 + o * nchmax * N, Q + o * nchmax * lqm::kQStride,
                           SS + o * nbmax, Vp + o * NP, Wp + o * NP, tau + o * N);
    };
    lqm_status st = LQM_OK;
    for (int p = 0; p < N - 1 && !st; p += lqm::kBnb) {
        const int pw = std::min(lqm::kBnb, N - 1 - p);
        {
            StageScope sc(c, ST_BTRD, Kc);
            for (int i = 0; i < pw; ++i)
                for (int h = 0; h < nh; ++h) col_step(h, p + i, i, 0);
        }
        if (hipGetLastError() != hipSuccess) {
            st = fail(LQM_ERR_DEVICE, "column-step launch failed");
            break;
        }
        // A22 -= V W^H + W V^H on rows/columns >= p + pw, tiles on or below the diagonal
        const int q = p + pw, n = N - q;
        for (int h = 0; h < nh && !st; ++h) {
            const size_t o = k0[h];
            lqm::GemmArgs<R> g = gemm_args<R>(n, n, pw, Vp + o * NP + q, N, (long long)NP, Wp + o * NP + q, N,
                                              (long long)NP, A + o * NN + (size_t)q * N + q, N, (long long)NN,
                                              {-1.0, 0.0}, {1.0, 0.0});
            g.npair = 2;
            g.A[1] = Wp + o * NP + q;
            g.B[1] = Vp + o * NP + q;
            g.lower = 1;
            c->stream = st_of[h];
            zgemm_gen<0, 1, R>(c, ST_BUPD, Kh[h], Kh[h], g);
            c->stream = s_main;
            if (hipGetLastError() != hipSuccess) st = fail(LQM_ERR_DEVICE, "trailing-update launch failed");
        }
    }
    if (!st) {
        StageScope sc(c, ST_BTRD, Kc);
        for (int h = 0; h < nh; ++h) col_step(h, N - 1, 0, 1);
    }
    if (split) {
        // join the side stream, also after a failed launch (nothing may stay in flight there)
        (void)hipEventRecord(c->ev_join, c->side);
        (void)hipStreamWaitEvent(s_main, c->ev_join, 0);
    }
    if (st) return st;
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

// One pass of the large-N FP64 path over the grid.  Outputs (any may be null): bands
// (ascending), evecs (N x N per k-point), rho on keys dRr (accumulated; zeroed by the
// caller) and eband.  occ: given occupations of the ascending eigenvalues (device).
lqm_status run_blocked(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                       int nRr, const int* dRr, cx<double>* rho, double* eband, double* bands,
                       double* evecs, const double* occ, const MuReq* mq) {
    lqm_context* c = m->ctx;
    const int N = m->N, d = m->d;
    if (N > lqm::kBMaxN) return fail(LQM_ERR_UNSUPPORTED, "blocked eigensolver supports N <= 4096");
    hipStream_t s = c->stream;
    const size_t NN = (size_t)N * N;
    const bool want_rho = rho != nullptr || eband != nullptr;
    const bool vectors = want_rho || evecs != nullptr;
    const int nchmax = std::max(1, (N - 1 + lqm::kHvCols - 1) / lqm::kHvCols);
    const int nbmax = lqm::hv_blocks(nchmax);
    const int nblk = std::max(1, (N - 1 + lqm::kBwy - 1) / lqm::kBwy);
    size_t perk = 16 * NN + 2 * 16 * (size_t)N * lqm::kBnb + 2 * 16 * (size_t)nchmax * N +
                  16 * (size_t)nchmax * lqm::kQStride + 8 * (size_t)nbmax + 64 * (size_t)N + 64;
    if (vectors) perk += 8 * NN + 24 * NN + 16 * NN + 16 * (size_t)nblk * lqm::kBwy * lqm::kBwy + 2 * 16 * 64 * (size_t)N;
    if (want_rho) perk += 16 * NN;
    const int Kmax = (int)std::max<size_t>(1, std::min<size_t>({std::min(kBlockedBudget, c->ws_cap) / perk, (size_t)Nk, (size_t)65535}));
    if (mq && Kmax < Nk) return kFallback;  // mu needs every eigenvalue before the first occupation
    begin_pass(c, LQM_PREC_FP64);
    const double wuni = 1.0 / (double)Nk;
    lqm_status st;
    const int RG = contract_group(nRr);
    const int groups = std::max(1, (nRr + RG - 1) / RG);
    const int tilesE = (int)((NN + 255) / 256);
    const int S = std::max(1, std::min(std::max(1, Kmax / kContractMinK), contract_splits(tilesE * groups)));
    cx<double>* part = nullptr;
    if (rho && nRr > 0 && (st = wsT(c, W_PART, (size_t)groups * S * RG * NN, &part))) return st;
    std::vector<double> hlam, hnorm;
    for (int k0 = 0; k0 < Nk; k0 += Kmax) {
        const int Kc = std::min(Kmax, Nk - k0);
        const double* ksc = ks + (size_t)k0 * d;
        const double* kwc = kw ? kw + k0 : nullptr;
        cx<double> *A, *Vp, *Wp, *P, *tau;
        double* vec;
        if ((st = wsT(c, W_BA, NN * Kmax, &A))) return st;
        if ((st = wsT(c, W_BVW, (size_t)2 * N * lqm::kBnb * Kmax, &Vp))) return st;
        Wp = Vp + (size_t)N * lqm::kBnb * Kmax;
        if ((st = wsT(c, W_BP, (size_t)2 * nchmax * N * Kmax + (size_t)nchmax * lqm::kQStride * Kmax + nbmax * Kmax,
                      &P)))
            return st;
        cx<double>* P2 = P + (size_t)nchmax * N * Kmax;
        cx<double>* Q = P2 + (size_t)nchmax * N * Kmax;
        double* SS = reinterpret_cast<double*>(Q + (size_t)nchmax * lqm::kQStride * Kmax);
        if ((st = wsT(c, W_BVEC, (size_t)(6 * N + 2) * Kmax, &vec))) return st;
        double* dd = vec;
        double* ee = dd + (size_t)N * Kmax;
        double* lamw = ee + (size_t)N * Kmax;
        double* fo = lamw + (size_t)N * Kmax;
        tau = reinterpret_cast<cx<double>*>(fo + (size_t)N * Kmax);
        double* tn = reinterpret_cast<double*>(tau + (size_t)N * Kmax);
        double* lam = bands ? bands + (size_t)k0 * N : lamw;
        // H(k), full Hermitian storage
        {
            StageScope sc(c, ST_BHK, Kc);
            if ((st = launch_hk_full(c, m, Kc, ksc, A))) return st;
        }
        LQM_HIP(hipGetLastError());
        if ((st = blocked_tridiag(c, N, Kc, A, Vp, Wp, P, P2, Q, SS, nchmax, nbmax, dd, ee, tau))) return st;
        {
            StageScope sc(c, ST_BEIG, Kc);
            const size_t shs = lqm::sturm_lds_bytes(N);
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(lqm::k_sturm_bisect),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shs);
            hipLaunchKernelGGL(lqm::k_sturm_bisect, dim3((N + 255) / 256, Kc), dim3(256), shs, s, N, dd, ee, lam, tn);
        }
        LQM_HIP(hipGetLastError());
        if (mq) {
            if ((st = mu_from_bands(c, N, Nk, lam, kw, mq->filling, T, &mu))) return st;
            *mq->mu = mu;
        }
        if (!vectors) continue;
        // columns needed: all for eigenvectors or given occupations, else the eigenvalues
        // whose occupation reaches kOccTiny (a prefix, f falling with the eigenvalue),
        // extended to the end of any cluster the cut would split
        int ncol = N;
        if (!evecs && !occ) {
            hlam.resize((size_t)Kc * N);
            hnorm.resize(Kc);
            LQM_HIP(hipMemcpyAsync(hlam.data(), lam, sizeof(double) * Kc * N, hipMemcpyDeviceToHost, s));
            LQM_HIP(hipMemcpyAsync(hnorm.data(), tn, sizeof(double) * Kc, hipMemcpyDeviceToHost, s));
            LQM_HIP(hipStreamSynchronize(s));
            ncol = 0;
            for (int k = 0; k < Kc; ++k) {
                int nk = 0;
                while (nk < N) {
                    const double x = hlam[(size_t)k * N + nk] - mu;
                    const double f = T > 0.0 ? 1.0 / (std::exp(x / T) + 1.0) : (x < 0.0 ? 1.0 : (x > 0.0 ? 0.0 : 0.5));
                    if (!(f >= kOccTiny)) break;
                    ++nk;
                }
                ncol = std::max(ncol, nk);
            }
            for (bool ext = ncol > 0; ext && ncol < N;) {
                ext = false;
                for (int k = 0; k < Kc && !ext; ++k)
                    ext = hlam[(size_t)k * N + ncol] - hlam[(size_t)k * N + ncol - 1] <= kClusterRel * hnorm[k];
                if (ext) ++ncol;
            }
        }
        double* eb = nullptr;
        if (want_rho) {
            if ((st = wsT(c, W_EB, (size_t)Kmax, &eb))) return st;
            StageScope sc(c, ST_REDUCE);
            hipLaunchKernelGGL(lqm::k_occ_eb, dim3(Kc), dim3(256), 0, s, N, lam, kwc, wuni, mu, T,
                               occ ? occ + (size_t)k0 * N : nullptr, fo, eb);
            if (eband) hipLaunchKernelGGL(lqm::k_sum, dim3(1), dim3(256), 0, s, Kc, eb, eband, 1);
            LQM_HIP(hipGetLastError());
        }
        if (ncol == 0) continue;  // every occupation below kOccTiny: rho_k = 0
        const int ldz = ncol;
        double *Zt, *scr;
        if ((st = wsT(c, W_BZ, (size_t)N * ldz * Kmax, &Zt))) return st;
        if ((st = wsT(c, W_BSC, (size_t)3 * N * ldz * Kmax, &scr))) return st;
        {
            StageScope sc(c, ST_BVEC, Kc);
            hipLaunchKernelGGL(lqm::k_invit, dim3((ncol + 63) / 64, Kc), dim3(64), 0, s, N, ncol, ldz, dd, ee, lam,
                               tn, Zt, scr, k0, c->d_status);
            hipLaunchKernelGGL(lqm::k_invit_orth, dim3((ncol + 3) / 4, Kc), dim3(256), 0, s, N, ncol, ldz, lam, tn,
                               kClusterRel, Zt);
        }
        LQM_HIP(hipGetLastError());
        // V = Q Z: Y explicit in A, T per block of kBwy reflectors, blocks applied last first
        cx<double>* V;
        if (evecs) {
            V = reinterpret_cast<cx<double>*>(evecs) + (size_t)k0 * NN;
        } else if ((st = wsT(c, W_BV, (size_t)N * ncol * Kmax, &V))) {
            return st;
        }
        cx<double> *Tw, *Wt;
        if ((st = wsT(c, W_BT, (size_t)nblk * lqm::kBwy * lqm::kBwy * Kmax, &Tw))) return st;
        if ((st = wsT(c, W_BWT, (size_t)2 * lqm::kBwy * ncol * Kmax, &Wt))) return st;
        cx<double>* TW = Wt + (size_t)lqm::kBwy * ncol * Kmax;
        {
            StageScope sc(c, ST_BBT, Kc);
            hipLaunchKernelGGL(lqm::k_make_y<double>, dim3((unsigned)((NN + 255) / 256), Kc), dim3(256), 0, s, N, A);
            if (N > 1)
                hipLaunchKernelGGL(lqm::k_wy_T<double>, dim3(nblk, Kc), dim3(256), 0, s, N, nblk, A, tau, Tw);
            hipLaunchKernelGGL(lqm::k_z_to_v<double>, dim3((N + 31) / 32, (ncol + 31) / 32, Kc), dim3(256), 0, s, N, ncol,
                               ldz, Zt, V);
        }
        LQM_HIP(hipGetLastError());
        const long long sV = (long long)N * ncol, sW = (long long)lqm::kBwy * ncol;
        for (int b = (N > 1 ? nblk - 1 : -1); b >= 0; --b) {
            const int jb = lqm::kBwy * b, nbb = std::min(lqm::kBwy, N - 1 - jb), r0 = jb + 1, Mr = N - r0;
            const cx<double>* Yb = A + (size_t)jb * N + r0;
            // Wt = Y^H V(r0:, :); TW = T Wt; V(r0:, :) -= Y TW
            zgemm_gen<1, 0>(c, ST_BBT, Kc, Kc,
                            gemm_args(nbb, ncol, Mr, Yb, N, (long long)NN, V + r0, N, sV, Wt, lqm::kBwy, sW,
                                      {1.0, 0.0}, {0.0, 0.0}));
            zgemm_gen<0, 0>(c, ST_BBT, Kc, Kc,
                            gemm_args(nbb, ncol, nbb, Tw + (size_t)b * lqm::kBwy * lqm::kBwy, lqm::kBwy,
                                      (long long)nblk * lqm::kBwy * lqm::kBwy, Wt, lqm::kBwy, sW, TW, lqm::kBwy, sW,
                                      {1.0, 0.0}, {0.0, 0.0}));
            zgemm_gen<0, 0>(c, ST_BBT, Kc, Kc,
                            gemm_args(Mr, ncol, nbb, Yb, N, (long long)NN, TW, lqm::kBwy, sW, V + r0, N, sV,
                                      {-1.0, 0.0}, {1.0, 0.0}));
        }
        LQM_HIP(hipGetLastError());
        if (!rho || nRr <= 0) continue;
        // rho_k = V diag(f) V^H, then the phase contraction onto the keys
        cx<double>* rhok;
        if ((st = wsT(c, W_RHOK, NN * Kmax, &rhok))) return st;
        {
            lqm::ZGemmArgs g = gemm_args(N, N, ncol, V, N, sV, V, N, sV, rhok, N, (long long)NN, {1.0, 0.0},
                                         {0.0, 0.0});
            g.ascale = fo;
            g.sS = N;
            g.lower = c->rr_pairs ? 2 : 0;  // the contraction reads rho_k on and above the diagonal
            zgemm_gen<0, 1>(c, ST_BRHO, Kc, Kc, g);
        }
        const int Sc = std::min(S, Kc);
        const int ksplit = (Kc + Sc - 1) / Sc;
        const int Sused = (Kc + ksplit - 1) / ksplit;
        if ((st = launch_contract(RG, c, N, Kc, d, ksc, kwc, wuni, nRr, dRr, rhok, ksplit, Sused, groups, part, rho))) return st;
        LQM_HIP(hipGetLastError());
    }
    return check_status(c, "blocked eigensolver");
}

// Split of the DOS reduction: S slices of esplit eigenvalues each (at most 1024 / tiles of 256
// frequencies), so the partials S x Nw hold at most max(262144, Nw) doubles.
struct DosSplit {
    int tilesO, S, esplit;
};
DosSplit dos_split(int Ne, int Nw) {
    DosSplit p;
    p.tilesO = (Nw + 255) / 256;
    p.S = std::max(1, std::min((Ne + 255) / 256, 1024 / p.tilesO));
    p.esplit = (Ne + p.S - 1) / p.S;
    p.S = (Ne + p.esplit - 1) / p.esplit;
    return p;
}

// k_dos over the slices into part [S x Nw], then the fixed-order fold into dos.
void dos_launch(lqm_context* c, int N, int Nk, const double* bands, const double* kw, int Nw,
                const double* om, double gamma, const DosSplit& p, double* part, double* dos) {
    const int Ne = N * Nk;
    hipLaunchKernelGGL(lqm::k_dos, dim3(p.tilesO, p.S), dim3(256), 0, c->stream, N, Ne, bands, kw, Nw, om,
                       gamma, p.esplit, part);
    const double scale = kw ? 1.0 : 1.0 / (double)Nk;
    hipLaunchKernelGGL(lqm::k_dos_fold, dim3(p.tilesO), dim3(256), 0, c->stream, Nw, p.S, part, scale,
                       dos);
}

void convert_launch_d2f(lqm_context* c, size_t n, const cx<double>* in, cx<float>* out) {
    const unsigned g = (unsigned)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL((lqm::k_convert_cx<double, float>), dim3(g), dim3(256), 0, c->stream, n, in, out);
}
void convert_launch_f2d(lqm_context* c, size_t n, const cx<float>* in, cx<double>* out) {
    const unsigned g = (unsigned)std::min<size_t>(8192, (n + 255) / 256);
    hipLaunchKernelGGL((lqm::k_convert_cx<float, double>), dim3(g), dim3(256), 0, c->stream, n, in, out);
}

// One pass of the large-N path in mixed precision (LQM_PREC_FP32_REFINED) over the grid: H(k)
// in FP64 (kept for the refinement) and rounded to FP32; the blocked Householder reduction,
// the compact-WY back-transformation and their trailing / panel products in FP32 (FP32
// matrix cores, FP32 accumulation: half the bytes of the HBM-bound matrix-vector products);
// the tridiagonal eigenpairs by FP64 bisection and inverse iteration of the FP32-rounded
// tridiagonal (all N vectors); then Ogita-Aishima sweeps in FP64 on the matrix cores over
// every k-point of the chunk: W = A X, S = X^H W, G = X^H X, the correction E with Rayleigh-
// Ritz on near-degenerate clusters (k_oa_correct, cluster rotations in global memory),
// X <- X + X E, until no k-point asks for another sweep (at most max(2, max_sweeps)).
// Eigenvalues are the refined Rayleigh quotients with their second-order correction; rho_k =
// X f X^H in FP64 over the columns whose occupation reaches kOccTiny.  Outputs as run_blocked
// (eigenvector output stays on the FP64 path).
constexpr size_t kBlockedMixedBudget = size_t(96) << 30;
lqm_status run_blocked_mixed(lqm_model* m, int Nk, const double* ks, const double* kw, double mu, double T,
                             int nRr, const int* dRr, cx<double>* rho, double* eband, double* bands,
                             const double* occ, const MuReq* mq) {
    lqm_context* c = m->ctx;
    const int N = m->N, d = m->d;
    if (N > lqm::kBMaxN) return fail(LQM_ERR_UNSUPPORTED, "blocked eigensolver supports N <= 4096");
    hipStream_t s = c->stream;
    const size_t NN = (size_t)N * N;
    const bool want_rho = rho != nullptr || eband != nullptr;
    const int nchmax = std::max(1, (N - 1 + lqm::kHvCols - 1) / lqm::kHvCols);
    const int nbmax = lqm::hv_blocks(nchmax);
    const int nblk = std::max(1, (N - 1 + lqm::kBwy - 1) / lqm::kBwy);
    const int maxclg = lqm::oa_max_clusters_global(N);
    size_t perk = 16 * NN + 8 * NN + 2 * 8 * (size_t)N * lqm::kBnb + 2 * 8 * (size_t)nchmax * N +
                  8 * (size_t)nchmax * lqm::kQStride + 8 * (size_t)nbmax + 64 * (size_t)N + 64 +
                  8 * NN + 24 * NN + 8 * NN + 8 * (size_t)nblk * lqm::kBwy * lqm::kBwy + 2 * 8 * 64 * (size_t)N +
                  4 * 16 * NN + 16 * (size_t)maxclg * lqm::kClusterMax * lqm::kClusterMax + 48;
    if (want_rho) perk += 16 * NN;
    const int Kmax = (int)std::max<size_t>(1, std::min<size_t>({std::min(kBlockedMixedBudget, c->ws_cap) / perk, (size_t)Nk, (size_t)65535}));
    if (mq && Kmax < Nk) return kFallback;  // mu needs every refined eigenvalue before the first occupation
    begin_pass(c, LQM_PREC_FP32_REFINED);
    const double wuni = 1.0 / (double)Nk;
    const cx<double> one = {1.0, 0.0}, zero = {0.0, 0.0};
    lqm_status st;
    const int RG = contract_group(nRr);
    const int groups = std::max(1, (nRr + RG - 1) / RG);
    const int tilesE = (int)((NN + 255) / 256);
    const int S = std::max(1, std::min(std::max(1, Kmax / kContractMinK), contract_splits(tilesE * groups)));
    cx<double>* part = nullptr;
    if (rho && nRr > 0 && (st = wsT(c, W_PART, (size_t)groups * S * RG * NN, &part))) return st;
    const int max_sw = std::max(2, c->refine_max);
    int sweeps_max = 0;
    double rmax = 0.0, dmax = 0.0, nnear = 0.0;
    std::vector<double> hstat, hlam;
    for (int k0 = 0; k0 < Nk; k0 += Kmax) {
        const int Kc = std::min(Kmax, Nk - k0);
        const double* ksc = ks + (size_t)k0 * d;
        const double* kwc = kw ? kw + k0 : nullptr;
        cx<double> *A, *X, *Wm, *Sm, *Gm, *Ug;
        cx<float> *A32, *Vp, *P, *X32, *Tw, *Wt;
        double *vec, *Zt, *scr, *lam, *stat;
        if ((st = wsT(c, W_BA, NN * Kmax, &A))) return st;
        if ((st = wsT(c, W_MA32, NN * Kmax, &A32))) return st;
        if ((st = wsT(c, W_BVW, (size_t)2 * N * lqm::kBnb * Kmax, &Vp))) return st;
        cx<float>* Wp = Vp + (size_t)N * lqm::kBnb * Kmax;
        if ((st = wsT(c, W_BP, (size_t)2 * nchmax * N * Kmax + (size_t)nchmax * lqm::kQStride * Kmax + nbmax * Kmax, &P)))
            return st;
        cx<float>* P2 = P + (size_t)nchmax * N * Kmax;
        cx<float>* Q = P2 + (size_t)nchmax * N * Kmax;
        double* SS = reinterpret_cast<double*>(Q + (size_t)nchmax * lqm::kQStride * Kmax);
        if ((st = wsT(c, W_BVEC, (size_t)(6 * N + 2) * Kmax, &vec))) return st;
        double* dd = vec;
        double* ee = dd + (size_t)N * Kmax;
        double* lamw = ee + (size_t)N * Kmax;  // eigenvalues of the FP32 tridiagonal
        double* fo = lamw + (size_t)N * Kmax;
        cx<float>* tau = reinterpret_cast<cx<float>*>(fo + (size_t)N * Kmax);
        double* tn = reinterpret_cast<double*>(tau) + (size_t)2 * N * Kmax;
        if ((st = wsT(c, W_BZ, NN * Kmax, &Zt))) return st;
        if ((st = wsT(c, W_BSC, 3 * NN * Kmax, &scr))) return st;
        if ((st = wsT(c, W_MX32, NN * Kmax, &X32))) return st;
        if ((st = wsT(c, W_BT, (size_t)nblk * lqm::kBwy * lqm::kBwy * Kmax, &Tw))) return st;
        if ((st = wsT(c, W_BWT, (size_t)2 * lqm::kBwy * N * Kmax, &Wt))) return st;
        cx<float>* TW = Wt + (size_t)lqm::kBwy * N * Kmax;
        if ((st = wsT(c, W_X, NN * Kmax, &X))) return st;
        if ((st = wsT(c, W_W, NN * Kmax, &Wm))) return st;
        if ((st = wsT(c, W_S, NN * Kmax, &Sm))) return st;
        if ((st = wsT(c, W_G, NN * Kmax, &Gm))) return st;
        if ((st = wsT(c, W_LAM, (size_t)N * Kmax, &lam))) return st;
        if ((st = wsT(c, W_STAT, (size_t)4 * Kmax, &stat))) return st;
        if ((st = wsT(c, W_MUG, (size_t)maxclg * lqm::kClusterMax * lqm::kClusterMax * Kmax, &Ug))) return st;
        // H(k) in FP64 (the refinement's A) and its FP32 rounding (the reduction's)
        {
            StageScope sc(c, ST_BHK, Kc);
            if ((st = launch_hk_full(c, m, Kc, ksc, A))) return st;
            convert_launch_d2f(c, NN * Kc, A, A32);
        }
        LQM_HIP(hipGetLastError());
        if ((st = blocked_tridiag<float>(c, N, Kc, A32, Vp, Wp, P, P2, Q, SS, nchmax, nbmax, dd, ee, tau))) return st;
        {
            StageScope sc(c, ST_BEIG, Kc);
            const size_t shs = lqm::sturm_lds_bytes(N);
            (void)hipFuncSetAttribute(reinterpret_cast<const void*>(lqm::k_sturm_bisect),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)shs);
            hipLaunchKernelGGL(lqm::k_sturm_bisect, dim3((N + 255) / 256, Kc), dim3(256), shs, s, N, dd, ee, lamw, tn);
        }
        {
            StageScope sc(c, ST_BVEC, Kc);
            hipLaunchKernelGGL(lqm::k_invit, dim3((N + 63) / 64, Kc), dim3(64), 0, s, N, N, N, dd, ee, lamw, tn, Zt,
                               scr, k0, c->d_status);
            hipLaunchKernelGGL(lqm::k_invit_orth, dim3((N + 3) / 4, Kc), dim3(256), 0, s, N, N, N, lamw, tn,
                               kClusterRel, Zt);
        }
        LQM_HIP(hipGetLastError());
        // X32 = Q Z in FP32: Y explicit in A32, T per block of kBwy reflectors, blocks last first
        {
            StageScope sc(c, ST_BBT, Kc);
            hipLaunchKernelGGL(lqm::k_make_y<float>, dim3((unsigned)((NN + 255) / 256), Kc), dim3(256), 0, s, N, A32);
            if (N > 1)
                hipLaunchKernelGGL(lqm::k_wy_T<float>, dim3(nblk, Kc), dim3(256), 0, s, N, nblk, A32, tau, Tw);
            hipLaunchKernelGGL(lqm::k_z_to_v<float>, dim3((N + 31) / 32, (N + 31) / 32, Kc), dim3(256), 0, s, N, N, N,
                               Zt, X32);
        }
        LQM_HIP(hipGetLastError());
        const long long sV = (long long)NN, sW = (long long)lqm::kBwy * N;
        for (int b = (N > 1 ? nblk - 1 : -1); b >= 0; --b) {
            const int jb = lqm::kBwy * b, nbb = std::min(lqm::kBwy, N - 1 - jb), r0 = jb + 1, Mr = N - r0;
            const cx<float>* Yb = A32 + (size_t)jb * N + r0;
            zgemm_gen<1, 0, float>(c, ST_BBT, Kc, Kc,
                                   gemm_args<float>(nbb, N, Mr, Yb, N, (long long)NN, X32 + r0, N, sV, Wt, lqm::kBwy,
                                                    sW, one, zero));
            zgemm_gen<0, 0, float>(c, ST_BBT, Kc, Kc,
                                   gemm_args<float>(nbb, N, nbb, Tw + (size_t)b * lqm::kBwy * lqm::kBwy, lqm::kBwy,
                                                    (long long)nblk * lqm::kBwy * lqm::kBwy, Wt, lqm::kBwy, sW, TW,
                                                    lqm::kBwy, sW, one, zero));
            zgemm_gen<0, 0, float>(c, ST_BBT, Kc, Kc,
                                   gemm_args<float>(Mr, N, nbb, Yb, N, (long long)NN, TW, lqm::kBwy, sW, X32 + r0, N,
                                                    sV, {-1.0, 0.0}, one));
        }
        LQM_HIP(hipGetLastError());
        convert_launch_f2d(c, NN * Kc, X32, X);
        LQM_HIP(hipGetLastError());
        // Ogita-Aishima sweeps: the first over every k-point of the chunk, later ones over the
        // k-points whose last correction asked for another (device index list)
        const size_t shc = lqm::oa_lds_bytes(N, true);
        allow_lds(lqm::k_oa_correct, shc);
        int sweeps = 0;
        hstat.resize((size_t)4 * Kc);
        std::vector<int> cur(Kc), next;
        for (int k = 0; k < Kc; ++k) cur[k] = k;
        const int* kidx = nullptr;
        int nsub = Kc;
        for (;;) {
            const int stg = sweeps == 0 ? ST_OA : ST_OA2;
            auto gg = [&](lqm::ZGemmArgs g) {
                g.kidx = kidx;
                return g;
            };
            zgemm_gen<0, 0>(c, stg, nsub, nsub, gg(gemm_args(N, N, N, A, N, (long long)NN, X, N, (long long)NN, Wm, N,
                                                             (long long)NN, one, zero)));  // W = A X
            // S and G are Hermitian: only their tiles on or below the diagonal are formed;
            // k_oa_correct completes them from the lower triangle
            auto lower = [](lqm::ZGemmArgs g) {
                g.lower = 1;
                return g;
            };
            zgemm_gen<1, 0>(c, stg, 0, nsub, lower(gg(gemm_args(N, N, N, X, N, (long long)NN, Wm, N, (long long)NN, Sm,
                                                                N, (long long)NN, one, zero))));  // S = X^H A X
            zgemm_gen<1, 0>(c, stg, 0, nsub, lower(gg(gemm_args(N, N, N, X, N, (long long)NN, X, N, (long long)NN, Gm,
                                                                N, (long long)NN, one, zero))));  // G = X^H X
            {
                StageScope sc(c, stg);
                hipLaunchKernelGGL(lqm::k_oa_correct, dim3(nsub), dim3(lqm::kOaT), shc, s, N, Sm, Gm, Wm, lam, stat,
                                   c->refine_thr, kidx, Ug);
                // X <- X + X E (E in W), formed in G
                if (!kidx) {
                    LQM_HIP(hipMemcpyAsync(Gm, X, sizeof(cx<double>) * NN * Kc, hipMemcpyDeviceToDevice, s));
                } else {
                    const int bx = (int)std::min<size_t>(1024, (NN + 255) / 256);
                    hipLaunchKernelGGL(lqm::k_copy_sub, dim3(bx, nsub), dim3(256), 0, s, (int)NN, nsub, kidx, X, Gm);
                }
            }
            zgemm_gen<0, 0>(c, stg, 0, nsub, gg(gemm_args(N, N, N, X, N, (long long)NN, Wm, N, (long long)NN, Gm, N,
                                                          (long long)NN, one, one)));
            if (!kidx) {
                std::swap(X, Gm);
            } else {
                const int bx = (int)std::min<size_t>(1024, (NN + 255) / 256);
                hipLaunchKernelGGL(lqm::k_copy_sub, dim3(bx, nsub), dim3(256), 0, s, (int)NN, nsub, kidx, Gm, X);
            }
            LQM_HIP(hipGetLastError());
            ++sweeps;
            LQM_HIP(hipMemcpyAsync(hstat.data(), stat, sizeof(double) * 4 * Kc, hipMemcpyDeviceToHost, s));
            LQM_HIP(hipStreamSynchronize(s));
            next.clear();
            int unresolved = 0;
            for (int k : cur) {
                rmax = std::max(rmax, hstat[4 * k]);
                dmax = std::max(dmax, hstat[4 * k + 1]);
                const int fl = (int)hstat[4 * k + 2];
                unresolved += (fl & 2) != 0;
                if (fl & 4) next.push_back(k);
                if (!std::isfinite(hstat[4 * k]) || !std::isfinite(hstat[4 * k + 3]))
                    return fail(LQM_ERR_NONFINITE, "mixed precision: non-finite refinement residual");
            }
            c->report[4] += (double)cur.size();
            if (next.empty()) break;
            if (sweeps >= max_sw) {
                nnear += unresolved;
                if (unresolved)
                    return fail(LQM_ERR_NOCONV, "mixed precision: a near-degenerate cluster wider than 8 "
                                                "eigenvalues could not be resolved; use the FP64 path");
                break;
            }
            int* dk;
            if ((st = wsT(c, W_KIDX, (size_t)Kc, &dk))) return st;
            LQM_HIP(hipMemcpyAsync(dk, next.data(), sizeof(int) * next.size(), hipMemcpyHostToDevice, s));
            kidx = dk;
            nsub = (int)next.size();
            cur.swap(next);
        }
        sweeps_max = std::max(sweeps_max, sweeps);
        // the workspace slots keep their pointers: X may now be the G slot's buffer
        double* bk = bands ? bands + (size_t)k0 * N : lamw;
        {
            StageScope sc(c, ST_REDUCE);
            hipLaunchKernelGGL(lqm::k_sort_evals, dim3(Kc), dim3(256), 0, s, N, lam, bk);
        }
        LQM_HIP(hipGetLastError());
        if (mq) {
            if ((st = mu_from_bands(c, N, Nk, bk, kw, mq->filling, T, &mu))) return st;
            *mq->mu = mu;
        }
        if (!want_rho) continue;
        double* eb;
        if ((st = wsT(c, W_EB, (size_t)Kmax, &eb))) return st;
        cx<double>* rhok = nullptr;
        if (rho && nRr > 0 && (st = wsT(c, W_RHOK, NN * Kmax, &rhok))) return st;
        const double* occk = occ ? occ + (size_t)k0 * N : nullptr;
        if (occk) {
            // given occupations by eigenvalue rank: Y = X diag(f) (into S), rho_k = Y X^H
            {
                StageScope sc(c, ST_REDUCE);
                hipLaunchKernelGGL(lqm::k_scale_occ, dim3(Kc), dim3(256), sizeof(double) * N, s, N, X, lam, kwc, wuni,
                                   mu, T, occk, Sm, eb, nullptr);
            }
            if (rhok) {
                lqm::ZGemmArgs g = gemm_args(N, N, N, Sm, N, (long long)NN, X, N, (long long)NN, rhok, N,
                                             (long long)NN, one, zero);
                g.lower = c->rr_pairs ? 2 : 0;  // the contraction reads rho_k on and above the diagonal
                zgemm_gen<0, 1>(c, ST_BRHO, Kc, Kc, g);
            }
        } else {
            {
                StageScope sc(c, ST_REDUCE);
                hipLaunchKernelGGL(lqm::k_occ_eb, dim3(Kc), dim3(256), 0, s, N, lam, kwc, wuni, mu, T, nullptr, fo, eb);
            }
            // columns up to the last whose occupation reaches kOccTiny (columns are in the
            // ascending order of the FP32 eigenvalues, so this is about the occupied half)
            int ncol = 0;
            if (rhok) {
                hlam.resize((size_t)Kc * N);
                LQM_HIP(hipMemcpyAsync(hlam.data(), lam, sizeof(double) * Kc * N, hipMemcpyDeviceToHost, s));
                LQM_HIP(hipStreamSynchronize(s));
                for (int k = 0; k < Kc; ++k)
                    for (int n = N - 1; n >= ncol; --n) {
                        const double x = hlam[(size_t)k * N + n] - mu;
                        const double f = T > 0.0 ? 1.0 / (std::exp(x / T) + 1.0) : (x < 0.0 ? 1.0 : (x > 0.0 ? 0.0 : 0.5));
                        if (!(f < kOccTiny)) {
                            ncol = n + 1;
                            break;
                        }
                    }
            }
            if (rhok && ncol > 0) {
                lqm::ZGemmArgs g = gemm_args(N, N, ncol, X, N, (long long)NN, X, N, (long long)NN, rhok, N,
                                             (long long)NN, one, zero);
                g.ascale = fo;
                g.sS = N;
                g.lower = c->rr_pairs ? 2 : 0;  // the contraction reads rho_k on and above the diagonal
                zgemm_gen<0, 1>(c, ST_BRHO, Kc, Kc, g);
            } else if (rhok) {
                LQM_HIP(hipMemsetAsync(rhok, 0, sizeof(cx<double>) * NN * Kc, s));
            }
        }
        if (eband) {
            StageScope sc(c, ST_REDUCE);
            hipLaunchKernelGGL(lqm::k_sum, dim3(1), dim3(256), 0, s, Kc, eb, eband, 1);
        }
        LQM_HIP(hipGetLastError());
        if (!rhok) continue;
        const int Sc = std::min(S, Kc);
        const int ksplit = (Kc + Sc - 1) / Sc;
        const int Sused = (Kc + ksplit - 1) / ksplit;
        if ((st = launch_contract(RG, c, N, Kc, d, ksc, kwc, wuni, nRr, dRr, rhok, ksplit, Sused, groups, part, rho)))
            return st;
        LQM_HIP(hipGetLastError());
    }
    c->report[0] = sweeps_max;
    c->report[1] = rmax;
    c->report[2] = dmax;
    c->report[3] = nnear;
    return check_status(c, "mixed-precision blocked eigensolver");
}

lqm_status dos_bands_dev(lqm_context* c, int N, int Nk, const double* bands, const double* kw,
                         int Nw, const double* om, double gamma, double* dos) {
    if (Nw <= 0) return LQM_OK;
    const DosSplit p = dos_split(N * Nk, Nw);
    double* part;
    lqm_status st = wsT(c, W_PART, (size_t)p.S * Nw + 1, &part);
    if (st) return st;
    StageScope sc(c, ST_DOS);
    dos_launch(c, N, Nk, bands, kw, Nw, om, gamma, p, part, dos);
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

uint64_t host_dkey(double v) {
    uint64_t u;
    std::memcpy(&u, &v, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
double host_dkey_inv(uint64_t k) {
    uint64_t u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

// r-th smallest (0-based) of n device doubles, by 8 byte-wise histogram passes.
lqm_status select_rank(lqm_context* c, int n, const double* x, long long r, double* out) {
    unsigned int* hist;
    lqm_status st = wsT(c, W_OUT, 256, &hist);
    if (st) return st;
    uint64_t prefix = 0, mask = 0;
    unsigned int h[256];
    const int G = std::max(1, std::min(1024, (n + 255) / 256));
    for (int shift = 56; shift >= 0; shift -= 8) {
        LQM_HIP(hipMemsetAsync(hist, 0, sizeof(unsigned int) * 256, c->stream));
        hipLaunchKernelGGL(lqm::k_radix_hist, dim3(G), dim3(256), 0, c->stream, n, x,
                           (unsigned long long)prefix, (unsigned long long)mask, shift, hist);
        LQM_HIP(hipMemcpyAsync(h, hist, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        LQM_HIP(hipStreamSynchronize(c->stream));
        int b = 0;
        for (; b < 256; ++b) {
            if (r < (long long)h[b]) break;
            r -= h[b];
        }
        if (b == 256) return fail(LQM_ERR_NONFINITE, "order statistic: rank out of range (NaN in bands?)");
        prefix |= (uint64_t)b << shift;
        mask |= (uint64_t)0xff << shift;
    }
    *out = host_dkey_inv(prefix);
    return LQM_OK;
}

lqm_status mu_from_bands(lqm_context* c, int N, int Nk, const double* bands, const double* kw,
                         double filling, double T, double* mu) {
    const long long L = (long long)N * Nk;
    if (L <= 0) return fail(LQM_ERR_ARG, "empty band set");
    if (T == 0.0) {
        // chemicalpotential_0! (Spectrum/chemicalpotential.jl:9-23)
        double a, b;
        lqm_status st;
        if (filling >= 1.0) return select_rank(c, (int)L, bands, L - 1, mu);
        if (filling <= 0.0) {
            if ((st = select_rank(c, (int)L, bands, 0, &a))) return st;
            *mu = a - 1.0;
            return LQM_OK;
        }
        long long i = 1 + (long long)std::floor(filling * (double)(L - 1));  // 1-based
        if ((st = select_rank(c, (int)L, bands, i - 1, &a))) return st;
        if ((st = select_rank(c, (int)L, bands, std::min(i, L - 1), &b))) return st;
        *mu = (a + b) / 2;
        return LQM_OK;
    }
    // T > 0: bracketing bisection of getfilling(mu) - filling on [min, max], on the device
    double *res, *ework = nullptr;
    lqm_status st = wsT(c, W_SMALL, 8, &res);
    if (st) return st;
    if (L > 16 * 1024 && (st = wsT(c, W_MUE, (size_t)L, &ework))) return st;
    {
        StageScope sc(c, ST_MU);
        hipLaunchKernelGGL(lqm::k_mu_bisect, dim3(1), dim3(1024), 0, c->stream, N, (int)L, bands, kw,
                           filling, T, res, ework);
    }
    LQM_HIP(hipGetLastError());
    double h[2];
    LQM_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    LQM_HIP(hipStreamSynchronize(c->stream));
    if (h[1] == 2.0) return fail(LQM_ERR_NONFINITE, "non-finite bands");
    if (h[1] == 1.0) return fail(LQM_ERR_ARG, "filling not bracketed by the band range");
    *mu = h[0];
    return LQM_OK;
}

// host <-> device staging helpers for the synchronous host-buffer entry points
lqm_status d2h(lqm_context* c, void* dst, const void* src, size_t bytes) {
    if (bytes) LQM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status stage_in(lqm_context* c, Slot s, const void* h, size_t bytes, void** dptr) {
    lqm_status st = ws(c, s, bytes, dptr);
    if (st) return st;
    if (bytes && h) LQM_HIP(hipMemcpyAsync(*dptr, h, bytes, hipMemcpyHostToDevice, c->stream));
    return LQM_OK;
}

lqm_status model_init(lqm_context* ctx, int N, int d, int nR, const int32_t* R, lqm_model** out) {
    if (!ctx || !out) return fail(LQM_ERR_ARG, "null context or output pointer");
    if (N <= 0 || d <= 0 || nR <= 0) return fail(LQM_ERR_ARG, "N, d and nR must be positive");
    if (N > 32767) return fail(LQM_ERR_UNSUPPORTED, "N above 32767");
    if (!R) return fail(LQM_ERR_ARG, "null key array");
    auto* m = new lqm_model();
    m->ctx = ctx;
    m->N = N;
    m->d = d;
    m->nR = nR;
    m->R.assign(R, R + (size_t)d * nR);
    hipError_t e = dev_malloc(&m->dR, sizeof(int) * d * nR);
    if (e == hipSuccess) e = dev_malloc(&m->dT, sizeof(cx<double>) * (size_t)N * N * nR);
    if (e != hipSuccess) {
        (void)dev_free(m->dR);
        delete m;
        return fail(LQM_ERR_DEVICE, std::string("model allocation: ") + hipGetErrorString(e));
    }
    *out = m;
    return LQM_OK;
}

// ---- Hartree-Fock: key bookkeeping, device state, the self-consistent loop ----------
constexpr int kMixBlocks = 256;

struct HFState {
    lqm_context* c = nullptr;
    int N = 0, d = 0, nRh = 0, nRv = 0, nRr = 0, nK = 0, qz = -1, rz = -1, flags = 3;
    std::vector<int32_t> RK;  // d x nK mean-field keys
    std::vector<int> map;     // per mean-field key: index in h, v, rho (-1 absent)
    std::vector<int> vr;      // per v key: index in rho (-1 absent)
    void* mem = nullptr;
    cx<double>*h = nullptr, *v = nullptr, *V0 = nullptr, *rho0 = nullptr, *rho1 = nullptr;
    int *dmap = nullptr, *dvr = nullptr;
    double* scal = nullptr;   // 16 doubles of per-iteration scalars
    double* part = nullptr;   // 2 x kMixBlocks mixing partials
    double* bands = nullptr;  // N x Nk eigenvalues of the mu pass
    lqm_model* hmf = nullptr;
    ~HFState() {
        if (c) (void)hipStreamSynchronize(c->stream);
        if (hmf) lqm_model_destroy(hmf);
        if (mem) (void)dev_free(mem);
    }
};

lqm_status hf_keys(int d, int n, const int32_t* R, const char* what, std::map<std::vector<int>, int>& out) {
    for (int r = 0; r < n; ++r) {
        std::vector<int> k(R + (size_t)r * d, R + (size_t)(r + 1) * d);
        if (!out.emplace(k, r).second) return fail(LQM_ERR_ARG, std::string("duplicate key in ") + what);
    }
    return LQM_OK;
}

lqm_status hf_setup(HFState& S, lqm_context* c, int N, int d, int nRh, const int32_t* Rh,
                    const double* Th, int nRv, const int32_t* Rv, const double* Tv, int nRr,
                    const int32_t* Rr, const double* rho, int flags, int Nk) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (N <= 0 || d <= 0) return fail(LQM_ERR_ARG, "N and d must be positive");
    if (nRh < 1 || !Rh || !Th) return fail(LQM_ERR_ARG, "h needs at least one key and its blocks");
    if (nRv < 0 || (nRv > 0 && (!Rv || !Tv))) return fail(LQM_ERR_ARG, "bad interaction v");
    if (nRr < 1 || !Rr || !rho) return fail(LQM_ERR_ARG, "rho needs at least one key and its blocks");
    if (flags < 0 || flags > 3) return fail(LQM_ERR_ARG, "flags must combine LQM_HF_HARTREE and LQM_HF_FOCK");
    S.c = c;
    S.N = N;
    S.d = d;
    S.nRh = nRh;
    S.nRv = nRv;
    S.nRr = nRr;
    S.flags = flags;
    std::map<std::vector<int>, int> kh, kv, kr;
    lqm_status st;
    if ((st = hf_keys(d, nRh, Rh, "h", kh)) || (st = hf_keys(d, nRv, Rv, "v", kv)) ||
        (st = hf_keys(d, nRr, Rr, "rho", kr)))
        return st;
    // H_MF keys: keys(h), then keys(v) \ keys(h) (initialize_hMF!, hartreefock.jl:186-201),
    // then the zero key for the Hartree shift
    std::vector<std::vector<int>> K;
    for (int r = 0; r < nRh; ++r) K.emplace_back(Rh + (size_t)r * d, Rh + (size_t)(r + 1) * d);
    for (int r = 0; r < nRv; ++r) {
        std::vector<int> k(Rv + (size_t)r * d, Rv + (size_t)(r + 1) * d);
        if (!kh.count(k)) K.push_back(k);
    }
    const std::vector<int> zero(d, 0);
    if (!kh.count(zero) && !kv.count(zero)) K.push_back(zero);
    S.nK = (int)K.size();
    S.RK.clear();
    S.map.assign(3 * S.nK, -1);
    for (int q = 0; q < S.nK; ++q) {
        S.RK.insert(S.RK.end(), K[q].begin(), K[q].end());
        auto a = kh.find(K[q]), b = kv.find(K[q]), e = kr.find(K[q]);
        if (a != kh.end()) S.map[3 * q] = a->second;
        if (b != kv.end()) S.map[3 * q + 1] = b->second;
        if (e != kr.end()) S.map[3 * q + 2] = e->second;
        if (K[q] == zero) S.qz = q;
    }
    auto z = kr.find(zero);
    S.rz = z == kr.end() ? -1 : z->second;
    if ((flags & 1) && S.rz < 0) return fail(LQM_ERR_ARG, "rho has no zero key (needed by the Hartree term)");
    S.vr.assign(std::max(nRv, 1), -1);
    for (int r = 0; r < nRv; ++r) {
        auto e = kr.find(std::vector<int>(Rv + (size_t)r * d, Rv + (size_t)(r + 1) * d));
        if (e != kr.end()) S.vr[r] = e->second;
    }
    // one device allocation: h, v, V0, rho0, rho1 | map, vr | scalars, partials, bands
    const size_t NN = (size_t)N * N;
    const size_t ncx = NN * ((size_t)nRh + std::max(nRv, 1) + 1 + 2 * (size_t)nRr);
    const size_t nint = 3 * (size_t)S.nK + S.vr.size();
    const size_t ndbl = 16 + 2 * kMixBlocks + (size_t)N * std::max(Nk, 0);
    const size_t bytes = ncx * sizeof(cx<double>) + ((nint * sizeof(int) + 255) / 256) * 256 + ndbl * sizeof(double);
    LQM_HIP(hipSetDevice(c->device));
    {
        hipError_t e = dev_malloc(&S.mem, bytes);
        if (e != hipSuccess) {
            S.mem = nullptr;
            return fail(LQM_ERR_DEVICE, std::string("Hartree-Fock state allocation: ") + hipGetErrorString(e));
        }
    }
    char* p = static_cast<char*>(S.mem);
    S.h = reinterpret_cast<cx<double>*>(p);
    S.v = S.h + NN * nRh;
    S.V0 = S.v + NN * std::max(nRv, 1);
    S.rho0 = S.V0 + NN;
    S.rho1 = S.rho0 + NN * nRr;
    p += ncx * sizeof(cx<double>);
    S.dmap = reinterpret_cast<int*>(p);
    S.dvr = S.dmap + 3 * S.nK;
    p += ((nint * sizeof(int) + 255) / 256) * 256;
    S.scal = reinterpret_cast<double*>(p);
    S.part = S.scal + 16;
    S.bands = S.part + 2 * kMixBlocks;
    hipStream_t s = c->stream;
    LQM_HIP(hipMemcpyAsync(S.h, Th, sizeof(cx<double>) * NN * nRh, hipMemcpyHostToDevice, s));
    if (nRv > 0) LQM_HIP(hipMemcpyAsync(S.v, Tv, sizeof(cx<double>) * NN * nRv, hipMemcpyHostToDevice, s));
    LQM_HIP(hipMemcpyAsync(S.rho0, rho, sizeof(cx<double>) * NN * nRr, hipMemcpyHostToDevice, s));
    LQM_HIP(hipMemcpyAsync(S.dmap, S.map.data(), sizeof(int) * S.map.size(), hipMemcpyHostToDevice, s));
    LQM_HIP(hipMemcpyAsync(S.dvr, S.vr.data(), sizeof(int) * S.vr.size(), hipMemcpyHostToDevice, s));
    LQM_HIP(hipMemsetAsync(S.scal, 0, sizeof(double) * 16, s));
    hipLaunchKernelGGL(lqm::k_blocksum, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, s, (int)NN,
                       nRv, S.v, S.V0);
    LQM_HIP(hipGetLastError());
    if ((st = model_init(c, N, d, S.nK, S.RK.data(), &S.hmf))) return st;
    LQM_HIP(hipMemcpyAsync(S.hmf->dR, S.RK.data(), sizeof(int) * S.RK.size(), hipMemcpyHostToDevice, s));
    return LQM_OK;
}

// H_MF[rho] into the mean-field model, eps_H / eps_F (+ imaginary parts) into scal[0..3].
lqm_status hf_apply(HFState& S, const cx<double>* rho) {
    lqm_context* c = S.c;
    const size_t tot = (size_t)S.N * S.N * S.nK;
    StageScope sc(c, ST_HF);
    hipLaunchKernelGGL(lqm::k_hf_build, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream,
                       S.N, S.nK, S.dmap, S.qz, S.rz, S.flags, S.h, S.v, S.V0, rho, S.hmf->dT);
    hipLaunchKernelGGL(lqm::k_hf_energy, dim3(1), dim3(256), 0, c->stream, S.N, S.nRv, S.dvr, S.rz,
                       S.flags, S.v, S.V0, rho, S.scal);
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

lqm_status hf_check_energy(const HFState& S, const double* sc) {
    const double tol = std::sqrt(2.220446049250313e-16) * std::max(1, S.N);
    if (!std::isfinite(sc[0]) || !std::isfinite(sc[1]))
        return fail(LQM_ERR_NONFINITE, "non-finite Hartree/Fock energy");
    if (std::fabs(sc[2]) > tol || std::fabs(sc[3]) > tol)
        return fail(LQM_ERR_ARG, "complex-valued Hartree/Fock energy: rho is not Hermitian");
    return LQM_OK;
}

lqm_status hf_outputs(HFState& S, int* nK_out, int32_t* RK_out, double* hmf_out) {
    if (nK_out) *nK_out = S.nK;
    if (RK_out) std::memcpy(RK_out, S.RK.data(), sizeof(int32_t) * S.RK.size());
    if (hmf_out)
        return d2h(S.c, hmf_out, S.hmf->dT, sizeof(cx<double>) * (size_t)S.N * S.N * S.nK);
    return LQM_OK;
}

// ---- BdG (Nambu) operators: the chemical potential enters through the zero key -------
// The model holds H_BdG at mu = 0 in Nambu order (electrons 0..d-1, holes d..2d-1, N = 2d).
// A pass at mu rewrites the zero-key block as base - mu tau_z from a saved copy of it (no
// drift across the many shifts of a root search) and the scope restores the copy on exit.
struct BdgScope {
    lqm_model* m = nullptr;
    int rz = -1;
    cx<double>* base = nullptr;
    ~BdgScope() {
        if (m && base && rz >= 0) {
            const size_t NN = (size_t)m->N * m->N;
            (void)hipMemcpyAsync(m->dT + (size_t)rz * NN, base, sizeof(cx<double>) * NN, hipMemcpyDeviceToDevice,
                                 m->ctx->stream);
            (void)hipStreamSynchronize(m->ctx->stream);
        }
    }
};

lqm_status bdg_begin(lqm_model* m, BdgScope& S) {
    if (m->N % 2) return fail(LQM_ERR_ARG, "a Nambu model has even dimension 2d");
    for (int r = 0; r < m->nR; ++r) {
        bool z = true;
        for (int i = 0; i < m->d; ++i) z = z && m->R[(size_t)r * m->d + i] == 0;
        if (z) {
            S.rz = r;
            break;
        }
    }
    if (S.rz < 0) return fail(LQM_ERR_ARG, "the Nambu model needs the zero key (chemical potential)");
    const size_t NN = (size_t)m->N * m->N;
    cx<double>* base;
    lqm_status st = wsT(m->ctx, W_BDGB, NN, &base);
    if (st) return st;
    LQM_HIP(hipMemcpyAsync(base, m->dT + (size_t)S.rz * NN, sizeof(cx<double>) * NN, hipMemcpyDeviceToDevice,
                           m->ctx->stream));
    S.m = m;
    S.base = base;
    return LQM_OK;
}

lqm_status bdg_shift(const BdgScope& S, double mu) {
    const size_t NN = (size_t)S.m->N * S.m->N;
    hipLaunchKernelGGL(lqm::k_bdg_shift, dim3((unsigned)((NN + 255) / 256)), dim3(256), 0, S.m->ctx->stream, S.m->N,
                       S.base, mu, S.m->dT + (size_t)S.rz * NN);
    LQM_HIP(hipGetLastError());
    return LQM_OK;
}

// Spectrum.filling(::BdGOperator) (Superconductivity/operators.jl:107-112): one full
// density-matrix pass of H - mu tau_z on the zero key, then Re tr(rho_el(0)) / d.
lqm_status bdg_filling(const BdgScope& S, int Nk, const double* ks, const double* kw, double mu, double T,
                       double* filling) {
    lqm_model* m = S.m;
    lqm_context* c = m->ctx;
    const size_t NN = (size_t)m->N * m->N;
    double* buf;
    lqm_status st = wsT(c, W_BDGR, 2 * NN + 4, &buf);
    if (st) return st;
    if ((st = bdg_shift(S, mu))) return st;
    std::vector<int32_t> R0(m->d, 0);
    if ((st = run_rho(m, Nk, ks, kw, 0.0, T, 1, R0.data(), buf, buf + 2 * NN, nullptr))) return st;
    hipLaunchKernelGGL(lqm::k_bdg_filling, dim3(1), dim3(256), 0, c->stream, m->N,
                       reinterpret_cast<const cx<double>*>(buf), buf + 2 * NN + 1);
    LQM_HIP(hipGetLastError());
    return d2h(c, filling, buf + 2 * NN + 1, sizeof(double));
}

// Roots.jl's bracketing midpoint for Float64: 0 across a sign change, else the midpoint of
// the IEEE bit patterns of |a| and |b| carrying the sign of a + b (at most 64 halvings).
double roots_middle(double a, double b) {
    if ((a < 0.0) != (b < 0.0) && a != 0.0 && b != 0.0) return 0.0;
    uint64_t ia, ib;
    const double aa = std::fabs(a), ab = std::fabs(b);
    std::memcpy(&ia, &aa, 8);
    std::memcpy(&ib, &ab, 8);
    const uint64_t im = (ia >> 1) + (ib >> 1) + (ia & ib & 1);
    double mid;
    std::memcpy(&mid, &im, 8);
    return (a + b) < 0.0 ? -mid : mid;
}

}  // namespace

extern "C" {

const char* lqm_last_error(void) { return g_err.c_str(); }
int lqm_abi_version(void) { return 1; }

lqm_status lqm_context_create(int device, int precision, lqm_context** out) {
    if (!out) return fail(LQM_ERR_ARG, "null output pointer");
    if (precision != LQM_PREC_FP64 && precision != LQM_PREC_FP32_REFINED)
        return fail(LQM_ERR_UNSUPPORTED, "unknown precision mode");
    int n = 0;
    LQM_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LQM_ERR_ARG, "device ordinal out of range");
    LQM_HIP(hipSetDevice(device));
    auto* c = new lqm_context();
    c->device = device;
    c->precision = precision;
    int lds = 65536;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess)
        c->lds_max = lds;
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming);
    if (e == hipSuccess) e = dev_malloc(&c->d_status, sizeof(int));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_status, 0, sizeof(int), c->own);
    if (e == hipSuccess) e = hipStreamSynchronize(c->own);
    if (e != hipSuccess) {
        if (c->own) (void)hipStreamDestroy(c->own);
        if (c->side) (void)hipStreamDestroy(c->side);
        if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
        if (c->ev_join) (void)hipEventDestroy(c->ev_join);
        delete c;
        return fail(LQM_ERR_DEVICE, std::string("context init: ") + hipGetErrorString(e));
    }
    lqm_status sst = scratch_address(device, &c->scratch);
    if (sst) {
        lqm_context_destroy(c);
        return sst;
    }
    c->stream = c->own;
    *out = c;
    return LQM_OK;
}

lqm_status lqm_context_destroy(lqm_context* c) {
    if (!c) return LQM_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    (void)hipStreamSynchronize(c->side);
    for (int s = 0; s < W_NSLOTS; ++s)
        if (c->buf[s]) (void)dev_free(c->buf[s]);
    (void)dev_free(c->d_status);
    if (c->pin) (void)hipHostFree(c->pin);
    (void)hipStreamDestroy(c->own);
    (void)hipStreamDestroy(c->side);
    (void)hipEventDestroy(c->ev_fork);
    (void)hipEventDestroy(c->ev_join);
    delete c;
    return LQM_OK;
}

lqm_status lqm_context_set_stream(lqm_context* c, void* s) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    c->stream = s ? static_cast<hipStream_t>(s) : c->own;
    return LQM_OK;
}

lqm_status lqm_context_set_progress(lqm_context* c, lqm_progress_fn fn, void* user) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    c->progress = fn;
    c->progress_user = user;
    return LQM_OK;
}

lqm_status lqm_context_set_graphs(lqm_context* c, int enable) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    c->graphs = enable ? 1 : 0;
    return LQM_OK;
}

lqm_status lqm_context_set_timing(lqm_context* c, int enable) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    c->timing = enable ? 1 : 0;
    return LQM_OK;
}

lqm_status lqm_context_stage_times(lqm_context* c, double* ms, int64_t* launches, int64_t* kpoints,
                                   int n) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    LQM_HIP(hipStreamSynchronize(c->stream));
    for (auto& t : c->pending) {
        float e = 0.f;
        if (hipEventElapsedTime(&e, t.a, t.b) == hipSuccess) {
            c->stage_ms[t.stage] += e;
            c->stage_n[t.stage] += 1;
            c->stage_k[t.stage] += t.kpts;
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    c->pending.clear();
    for (int i = 0; i < n && i < ST_N; ++i) {
        if (ms) ms[i] = c->stage_ms[i];
        if (launches) launches[i] = c->stage_n[i];
        if (kpoints) kpoints[i] = c->stage_k[i];
        c->stage_ms[i] = 0;
        c->stage_n[i] = 0;
        c->stage_k[i] = 0;
    }
    return LQM_OK;
}

lqm_status lqm_context_set_refinement(lqm_context* c, int max_sweeps) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (max_sweeps < 1 || max_sweeps > 8) return fail(LQM_ERR_ARG, "max_sweeps must be in 1..8");
    c->refine_max = max_sweeps;
    return LQM_OK;
}

lqm_status lqm_context_set_refinement_threshold(lqm_context* c, double thr) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (!(thr < 1.0)) return fail(LQM_ERR_ARG, "threshold must be below 1");
    c->refine_thr = thr;
    return LQM_OK;
}

lqm_status lqm_context_refinement_report(lqm_context* c, double* out) {
    if (!c || !out) return fail(LQM_ERR_ARG, "null argument");
    for (int i = 0; i < 6; ++i) out[i] = c->report[i];
    return LQM_OK;
}

namespace {
// Largest variant index of a kind (the range lqm_context_set_variant accepts).
int variant_max(int kind) { return kind == VK_TRIDIAG ? 4 : kind == VK_REFINE ? 2 : kind == VK_QL ? 3 : 1; }
std::string device_arch(int device) {
    hipDeviceProp_t p{};
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return "unknown";
    std::string a(p.gcnArchName);
    return a.substr(0, a.find(':'));  // gfx950, without the feature suffixes
}
}  // namespace

lqm_status lqm_context_set_variant(lqm_context* c, int kind, int variant) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (kind < 0 || kind >= VK_N || variant < -1 || variant > variant_max(kind)) return fail(LQM_ERR_ARG, "bad variant");
    c->forced[kind] = variant;
    return LQM_OK;
}


lqm_status lqm_context_save_tuning(lqm_context* c, const char* path) {
    if (!c || !path) return fail(LQM_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "w");
    if (!f) return fail(LQM_ERR_ARG, std::string("cannot write ") + path);
    std::fprintf(f, "lqm-tuning 1 %s\n", device_arch(c->device).c_str());
    for (const auto& kv : c->tuned) std::fprintf(f, "%d %d %d\n", kv.first >> 20, kv.first & 0xFFFFF, kv.second);
    const bool ok = std::fclose(f) == 0;
    return ok ? LQM_OK : fail(LQM_ERR_ARG, std::string("cannot write ") + path);
}

lqm_status lqm_context_load_tuning(lqm_context* c, const char* path) {
    if (!c || !path) return fail(LQM_ERR_ARG, "null argument");
    FILE* f = std::fopen(path, "r");
    if (!f) return fail(LQM_ERR_ARG, std::string("cannot read ") + path);
    char arch[64] = {0};
    int version = 0;
    std::map<int, int> got;
    lqm_status st = LQM_OK;
    if (std::fscanf(f, "lqm-tuning %d %63s", &version, arch) != 2 || version != 1) {
        st = fail(LQM_ERR_ARG, "not a tuning file (version 1)");
    } else if (device_arch(c->device) != arch) {
        st = fail(LQM_ERR_ARG, std::string("tuning file is for ") + arch + ", this device is " + device_arch(c->device));
    } else {
        int kind, N, v, r;
        while ((r = std::fscanf(f, "%d %d %d", &kind, &N, &v)) == 3) {
            if (kind < 0 || kind >= VK_N || N < 1 || N > 0xFFFFF || v < 0 || v > variant_max(kind)) {
                st = fail(LQM_ERR_ARG, "bad entry in the tuning file");
                break;
            }
            got[(kind << 20) | N] = v;
        }
        if (st == LQM_OK && r != EOF) st = fail(LQM_ERR_ARG, "unreadable entry in the tuning file");
    }
    std::fclose(f);
    if (st) return st;
    for (const auto& kv : got) c->tuned[kv.first] = kv.second;  // merged over what is there
    return LQM_OK;
}

lqm_status lqm_context_last_precision(lqm_context* c, int* precision) {
    if (!c || !precision) return fail(LQM_ERR_ARG, "null argument");
    *precision = c->last_prec;
    return LQM_OK;
}

lqm_status lqm_context_get_variant(lqm_context* c, int kind, int N, int* variant) {
    if (!c || !variant) return fail(LQM_ERR_ARG, "null argument");
    if (kind < 0 || kind >= VK_N) return fail(LQM_ERR_ARG, "bad variant kind");
    if (c->forced[kind] >= 0) {
        *variant = c->forced[kind];
        return LQM_OK;
    }
    auto it = c->tuned.find((kind << 20) | N);
    *variant = it == c->tuned.end() ? -1 : it->second;
    return LQM_OK;
}

const char* lqm_stage_name(int stage) { return (stage >= 0 && stage < ST_N) ? kStageNames[stage] : ""; }

lqm_status lqm_context_synchronize(lqm_context* c) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    LQM_HIP(hipStreamSynchronize(c->stream));
    LQM_HIP(hipStreamSynchronize(c->side));  // joined into the stream already; idle here
    return LQM_OK;
}

lqm_status lqm_model_create(lqm_context* ctx, int N, int d, int nR, const int32_t* R,
                            const double* T, lqm_model** out) {
    if (!T) return fail(LQM_ERR_ARG, "null hopping array");
    lqm_status st = model_init(ctx, N, d, nR, R, out);
    if (st) return st;
    lqm_model* m = *out;
    LQM_HIP(hipMemcpyAsync(m->dR, R, sizeof(int) * d * nR, hipMemcpyHostToDevice, ctx->stream));
    LQM_HIP(hipMemcpyAsync(m->dT, T, sizeof(cx<double>) * (size_t)N * N * nR, hipMemcpyHostToDevice,
                           ctx->stream));
    LQM_HIP(hipStreamSynchronize(ctx->stream));
    return LQM_OK;
}

lqm_status lqm_model_create_dev(lqm_context* ctx, int N, int d, int nR, const int32_t* R,
                                const double* T_dev, lqm_model** out) {
    if (!T_dev) return fail(LQM_ERR_ARG, "null hopping array");
    lqm_status st = model_init(ctx, N, d, nR, R, out);
    if (st) return st;
    lqm_model* m = *out;
    LQM_HIP(hipMemcpyAsync(m->dR, R, sizeof(int) * d * nR, hipMemcpyHostToDevice, ctx->stream));
    LQM_HIP(hipMemcpyAsync(m->dT, T_dev, sizeof(cx<double>) * (size_t)N * N * nR,
                           hipMemcpyDeviceToDevice, ctx->stream));
    LQM_HIP(hipStreamSynchronize(ctx->stream));
    return LQM_OK;
}

lqm_status lqm_model_destroy(lqm_model* m) {
    if (!m) return LQM_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    (void)dev_free(m->dR);
    (void)dev_free(m->dT);
    delete m;
    return LQM_OK;
}

lqm_status lqm_fouriersum_dev(lqm_model* m, int Nk, const double* ks_dev, double* Hk_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0) return Nk == 0 ? LQM_OK : fail(LQM_ERR_ARG, "Nk < 0");
    if (!ks_dev || !Hk_dev) return fail(LQM_ERR_ARG, "null buffer");
    const size_t sh = sizeof(cx<double>) * m->nR;
    auto kH = lqm::k_fouriersum;
    allow_lds(kH, sh);
    hipLaunchKernelGGL(kH, dim3(Nk), dim3(256), sh, m->ctx->stream, m->N, m->d, m->nR, m->dR, m->dT,
                       ks_dev, reinterpret_cast<cx<double>*>(Hk_dev));
    LQM_HIP(hipGetLastError());
    LQM_HIP(hipStreamSynchronize(m->ctx->stream));
    return LQM_OK;
}

lqm_status lqm_fouriersum(lqm_model* m, int Nk, const double* ks, double* Hk) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0) return Nk == 0 ? LQM_OK : fail(LQM_ERR_ARG, "Nk < 0");
    lqm_context* c = m->ctx;
    void *dks, *dH;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if ((st = ws(c, W_OUT, sizeof(cx<double>) * (size_t)m->N * m->N * Nk, &dH))) return st;
    if ((st = lqm_fouriersum_dev(m, Nk, (const double*)dks, (double*)dH))) return st;
    return d2h(c, Hk, dH, sizeof(cx<double>) * (size_t)m->N * m->N * Nk);
}

lqm_status lqm_bandmatrix_dev(lqm_model* m, int Nk, const double* ks_dev, double* bands_dev,
                              double* evecs_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    return run_bands(m, Nk, ks_dev, bands_dev, evecs_dev);
}

lqm_status lqm_bandmatrix(lqm_model* m, int Nk, const double* ks, double* bands, double* evecs) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0) return Nk == 0 ? LQM_OK : fail(LQM_ERR_ARG, "Nk < 0");
    lqm_context* c = m->ctx;
    const size_t N = m->N;
    void *dks, *db, *dv = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if ((st = ws(c, W_BANDS, sizeof(double) * N * Nk, &db))) return st;
    if (evecs && (st = ws(c, W_OUT, sizeof(cx<double>) * N * N * Nk, &dv))) return st;
    if ((st = run_bands(m, Nk, (const double*)dks, (double*)db, (double*)dv))) return st;
    if ((st = d2h(c, bands, db, sizeof(double) * N * Nk))) return st;
    if (evecs) return d2h(c, evecs, dv, sizeof(cx<double>) * N * N * Nk);
    return LQM_OK;
}

lqm_status lqm_dos_from_bands_dev(lqm_context* c, int N, int Nk, const double* bands_dev,
                                  const double* kweights_dev, int Nw, const double* omegas_dev,
                                  double gamma, double* dos_dev) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (N <= 0 || Nk <= 0 || Nw < 0) return fail(LQM_ERR_ARG, "bad sizes");
    lqm_status st = dos_bands_dev(c, N, Nk, bands_dev, kweights_dev, Nw, omegas_dev, gamma, dos_dev);
    if (st) return st;
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_dos_from_bands_dev_noalloc(lqm_context* c, int N, int Nk, const double* bands_dev,
                                          const double* kweights_dev, int Nw, const double* omegas_dev,
                                          double gamma, double* dos_dev) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (N <= 0 || Nk <= 0 || Nw < 0) return fail(LQM_ERR_ARG, "bad sizes");
    if ((long long)N * Nk > INT32_MAX) return fail(LQM_ERR_UNSUPPORTED, "N * Nk above 2^31 - 1");
    if ((size_t)Nw > kScratchDoubles)
        return fail(LQM_ERR_UNSUPPORTED, "Nw above " + std::to_string(kScratchDoubles) +
                                             ", the frequencies the scratch partials hold");
    if (!c->scratch) return fail(LQM_ERR_DEVICE, "scratch address not resolved for this device");
    if (Nw == 0) return LQM_OK;
    if (!bands_dev || !omegas_dev || !dos_dev) return fail(LQM_ERR_ARG, "null buffer");
    const DosSplit p = dos_split(N * Nk, Nw);
    if ((size_t)p.S * Nw > kScratchDoubles) return fail(LQM_ERR_UNSUPPORTED, "DOS partials exceed the scratch");
    dos_launch(c, N, Nk, bands_dev, kweights_dev, Nw, omegas_dev, gamma, p, c->scratch, dos_dev);
    LQM_HIP(hipGetLastError());
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_dos(lqm_model* m, int Nk, const double* ks, const double* kweights, int Nw,
                   const double* omegas, double gamma, double* dos) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || Nw < 0) return fail(LQM_ERR_ARG, "bad sizes");
    lqm_context* c = m->ctx;
    const size_t N = m->N;
    void *dks, *dkw = nullptr, *dom, *db, *dd;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = stage_in(c, W_OM, omegas, sizeof(double) * std::max(Nw, 1), &dom))) return st;
    if ((st = ws(c, W_BANDS, sizeof(double) * N * Nk, &db))) return st;
    if ((st = ws(c, W_OUT, sizeof(double) * std::max(Nw, 1), &dd))) return st;
    if ((st = run_bands(m, Nk, (const double*)dks, (double*)db, nullptr))) return st;
    if ((st = dos_bands_dev(c, (int)N, Nk, (double*)db, (double*)dkw, Nw, (double*)dom, gamma,
                            (double*)dd)))
        return st;
    return d2h(c, dos, dd, sizeof(double) * Nw);
}

lqm_status lqm_chemicalpotential_from_bands_dev(lqm_context* c, int N, int Nk, const double* bands_dev,
                                                const double* kweights_dev, double filling, double T,
                                                double* mu) {
    if (!c || !mu || !bands_dev) return fail(LQM_ERR_ARG, "null argument");
    if (N <= 0 || Nk <= 0) return fail(LQM_ERR_ARG, "bad sizes");
    if (T < 0) return fail(LQM_ERR_ARG, "negative temperature");
    return mu_from_bands(c, N, Nk, bands_dev, kweights_dev, filling, T, mu);
}

lqm_status lqm_chemicalpotential(lqm_model* m, int Nk, const double* ks, const double* kweights,
                                 double filling, double T, double* mu) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0) return fail(LQM_ERR_ARG, "Nk must be positive");
    lqm_context* c = m->ctx;
    void *dks, *dkw = nullptr, *db;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = ws(c, W_BANDS, sizeof(double) * m->N * Nk, &db))) return st;
    if ((st = run_bands(m, Nk, (const double*)dks, (double*)db, nullptr))) return st;
    return lqm_chemicalpotential_from_bands_dev(c, m->N, Nk, (double*)db, (double*)dkw, filling, T, mu);
}

lqm_status lqm_autotune(lqm_model* m, int Nk, const double* ks_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks_dev) return fail(LQM_ERR_ARG, "autotune needs a non-empty device k sample");
    lqm_context* c = m->ctx;
    const int N = m->N;
    int saved[VK_N];
    for (int kind = 0; kind < VK_N; ++kind) saved[kind] = c->forced[kind];
    const int saved_timing = c->timing;
    c->timing = 0;
    double* rho;
    if ((st = wsT(c, W_OUT, (size_t)2 * N * N + 2, &rho))) return st;
    const int32_t R0[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int nr = m->d <= 8 ? 1 : 0;
    hipEvent_t a, b;
    LQM_HIP(hipEventCreate(&a));
    LQM_HIP(hipEventCreate(&b));
    // one kind at a time, the other left at its current choice
    for (int kind = 0; kind < VK_N; ++kind) {
        if (kind == VK_REFINE) continue;  // a mixed-path choice with a fixed default
        float best = 1e30f;
        int bestv = 0;
        const int vmax = kind == VK_TRIDIAG ? (N <= 64 ? 4 : 1) : kind == VK_QL ? 3 : 1;
        if (kind == VK_BT && N > 64) continue;
        if (kind == VK_BLOCKED && (N <= 64 || N > 512)) continue;  // per-k kernels / large-N path only
        if (kind == VK_TRIEIG && (N <= 16 || N > 64)) continue;    // feeds the compact-WY kernel only
        for (int v = 0; v <= vmax; ++v) {
            c->forced[kind] = v;
            float t = 0.f;
            for (int rep = 0; rep < 3 && !st; ++rep) {
                LQM_HIP(hipEventRecord(a, c->stream));
                st = run_rho(m, Nk, ks_dev, nullptr, 0.0, 0.01, nr, R0, rho, rho + 2 * N * N, nullptr);
                LQM_HIP(hipEventRecord(b, c->stream));
                LQM_HIP(hipEventSynchronize(b));
                float e = 0.f;
                LQM_HIP(hipEventElapsedTime(&e, a, b));
                if (rep > 0) t += e;  // first run warms up
            }
            if (st) break;
            if (t < best) {
                best = t;
                bestv = v;
            }
        }
        if (st) break;
        c->tuned[(kind << 20) | N] = bestv;
        c->forced[kind] = saved[kind];
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    for (int kind = 0; kind < VK_N; ++kind) c->forced[kind] = saved[kind];
    c->timing = saved_timing;
    return st;
}

lqm_status lqm_densitymatrix_dev(lqm_model* m, int Nk, const double* ks_dev,
                                 const double* kweights_dev, double mu, double T, int nRr,
                                 const int32_t* Rr, double* rho_dev, double* eband_dev,
                                 double* bands_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    return run_rho(m, Nk, ks_dev, kweights_dev, mu, T, nRr, Rr, rho_dev, eband_dev, bands_dev);
}

lqm_status lqm_densitymatrix(lqm_model* m, int Nk, const double* ks, const double* kweights,
                             double mu, double T, int nRr, const int32_t* Rr, double* rho,
                             double* eband) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (!rho && nRr > 0) return fail(LQM_ERR_ARG, "null rho");
    lqm_context* c = m->ctx;
    const size_t NN = (size_t)m->N * m->N;
    void *dks, *dkw = nullptr, *drho, *deb;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = ws(c, W_OUT, sizeof(cx<double>) * NN * std::max(nRr, 1) + 16, &drho))) return st;
    deb = static_cast<char*>(drho) + sizeof(cx<double>) * NN * std::max(nRr, 1);
    if ((st = run_rho(m, Nk, (const double*)dks, (const double*)dkw, mu, T, nRr, Rr, (double*)drho,
                      (double*)deb, nullptr)))
        return st;
    if (nRr > 0 && (st = d2h(c, rho, drho, sizeof(cx<double>) * NN * nRr))) return st;
    double h = 0;
    if ((st = d2h(c, &h, deb, sizeof(double)))) return st;
    if (eband) *eband = h;
    return LQM_OK;
}

lqm_status lqm_hartreefock_meanfield(lqm_context* ctx, int N, int d, int nRh, const int32_t* Rh,
                                     const double* Th, int nRv, const int32_t* Rv, const double* Tv,
                                     int nRr, const int32_t* Rr, const double* rho, int flags,
                                     int* nK_out, int32_t* RK_out, double* hmf_out,
                                     double* energies_out) {
    HFState S;
    lqm_status st = hf_setup(S, ctx, N, d, nRh, Rh, Th, nRv, Rv, Tv, nRr, Rr, rho, flags, 0);
    if (st) return st;
    if ((st = hf_apply(S, S.rho0))) return st;
    double sc[4];
    if ((st = d2h(ctx, sc, S.scal, sizeof(sc)))) return st;
    if ((st = hf_check_energy(S, sc))) return st;
    if (energies_out) {
        energies_out[0] = sc[0];
        energies_out[1] = sc[1];
    }
    return hf_outputs(S, nK_out, RK_out, hmf_out);
}

lqm_status lqm_hartreefock_scf(lqm_context* ctx, int N, int d, int nRh, const int32_t* Rh,
                               const double* Th, int nRv, const int32_t* Rv, const double* Tv,
                               int nRr, const int32_t* Rr, double* rho, int Nk, const double* ks,
                               const double* kweights, double filling, double T, double beta,
                               double tol, int max_iter, int flags, double* result, int* converged,
                               double* history, int* nK_out, int32_t* RK_out, double* hmf_out) {
    if (Nk <= 0 || !ks) return fail(LQM_ERR_ARG, "the SCF needs a non-empty k-grid");
    if (max_iter < 1) return fail(LQM_ERR_ARG, "max_iter must be positive");
    if (!(T >= 0.0)) return fail(LQM_ERR_ARG, "negative temperature");
    if (!std::isfinite(beta) || !std::isfinite(tol)) return fail(LQM_ERR_ARG, "non-finite beta or tol");
    if (!result || !converged) return fail(LQM_ERR_ARG, "null result or converged pointer");
    HFState S;
    lqm_status st = hf_setup(S, ctx, N, d, nRh, Rh, Th, nRv, Rv, Tv, nRr, Rr, rho, flags, Nk);
    if (st) return st;
    lqm_context* c = ctx;
    void *dks, *dkw = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    const double* ksd = static_cast<const double*>(dks);
    const double* kwd = static_cast<const double*>(dkw);
    const size_t nrho = (size_t)N * N * nRr;
    double sc[10] = {};
    double mu = 0.0, eband = 0.0, residual = INFINITY;
    int it = 0;
    bool conv = false;
    int hstat = 0;
    // Graph-driven loop: the per-k FP64 kernels at T > 0 with the grid in one chunk keep mu on
    // the device, so an iteration needs no host round trip.  A device step kernel does the
    // host loop's bookkeeping (energy, residual, history row, stop test); iterations are
    // captured once in a HIP graph of kScfGraphIters iterations and replayed until the control
    // block says done (iterations after that recompute the frozen state and change nothing).
    const bool graph_mode = c->graphs && !c->timing && T > 0.0 && c->precision == LQM_PREC_FP64 &&
                            !use_blocked(c, N) && make_plan(c, N, S.nK, Nk, true, true).Kc >= Nk;
    if (graph_mode) {
        constexpr int kScfGraphIters = 8;
        double* ctl;
        const size_t nctl = 8, nh = (size_t)3 * max_iter;
        if ((st = wsT(c, W_SCF, nctl + nh + ((size_t)nRr * d + 1) / 2 + 1, &ctl))) return st;
        double* dhist = ctl + nctl;
        int* dRr = reinterpret_cast<int*>(dhist + nh);
        zero_dev(c, ctl, sizeof(double) * nctl);
        LQM_HIP(hipMemcpyAsync(dRr, Rr, sizeof(int) * nRr * d, hipMemcpyHostToDevice, c->stream));
        const double imag_tol = std::sqrt(2.220446049250313e-16) * std::max(1, N);
        auto iteration = [&]() -> lqm_status {
            lqm_status s2;
            if ((s2 = hf_apply(S, S.rho0))) return s2;
            const MuReq mq{filling, &mu, S.scal + 8};
            s2 = run_rho(S.hmf, Nk, ksd, kwd, 0.0, T, nRr, Rr, reinterpret_cast<double*>(S.rho1), S.scal + 4,
                         S.bands, nullptr, &mq, dRr);
            if (s2) return s2;
            if (!mq.on_device) return fail(LQM_ERR_DEVICE, "SCF graph: the pass did not keep mu on the device");
            hipLaunchKernelGGL(lqm::k_mix, dim3(kMixBlocks), dim3(256), 0, c->stream, nrho, beta, S.rho1, S.rho0,
                               S.part, ctl);
            hipLaunchKernelGGL(lqm::k_fold2, dim3(1), dim3(256), 0, c->stream, kMixBlocks, S.part, S.scal + 5);
            hipLaunchKernelGGL(lqm::k_scf_step, dim3(1), dim3(64), 0, c->stream, S.scal, history ? dhist : nullptr,
                               ctl, tol, max_iter, imag_tol, 1);
            LQM_HIP(hipGetLastError());
            return LQM_OK;
        };
        double hc[8] = {};
        auto poll = [&]() -> lqm_status {
            LQM_HIP(hipMemcpyAsync(hc, ctl, sizeof(hc), hipMemcpyDeviceToHost, c->stream));
            LQM_HIP(hipMemcpyAsync(&hstat, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            LQM_HIP(hipStreamSynchronize(c->stream));
            if (hstat) return check_status(c, "densitymatrix");
            return LQM_OK;
        };
        // the first iteration eagerly (it also sizes every workspace the captured ones use)
        if ((st = iteration()) || (st = poll())) return st;
        if (hc[0] == 0.0) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            lqm_status cst = LQM_OK;
            hipError_t ee;
            {
                std::lock_guard<std::mutex> lk(g_cap_mu);
                LQM_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
                c->capturing = true;
                for (int g = 0; g < kScfGraphIters && !cst; ++g) cst = iteration();
                c->capturing = false;
                ee = hipStreamEndCapture(c->stream, &graph);
            }
            if (cst) {
                if (graph) (void)hipGraphDestroy(graph);
                return cst;
            }
            if (ee != hipSuccess) return fail(LQM_ERR_DEVICE, std::string("SCF graph capture: ") + hipGetErrorString(ee));
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(graph);
                return fail(LQM_ERR_DEVICE, std::string("SCF graph instantiation: ") + hipGetErrorString(ie));
            }
            while (hc[0] == 0.0 && !st) {
                const hipError_t le = hipGraphLaunch(exec, c->stream);
                if (le != hipSuccess) st = fail(LQM_ERR_DEVICE, std::string("SCF graph launch: ") + hipGetErrorString(le));
                if (!st) st = poll();
            }
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            if (st) return st;
        }
        it = (int)hc[1];
        residual = hc[2];
        conv = hc[4] != 0.0;
        switch ((int)hc[3]) {
            case 1: return fail(LQM_ERR_NONFINITE, "non-finite bands");
            case 2: return fail(LQM_ERR_ARG, "filling not bracketed by the band range");
            case 3: return fail(LQM_ERR_NONFINITE, "non-finite Hartree/Fock energy");
            case 4: return fail(LQM_ERR_ARG, "complex-valued Hartree/Fock energy: rho is not Hermitian");
            case 5: return fail(LQM_ERR_NONFINITE, "non-finite SCF residual");
            default: break;
        }
        eband = hc[5];
        mu = hc[6];
        if (history && it > 0 && (st = d2h(c, history, dhist, sizeof(double) * 3 * it))) return st;
    }
    while (!graph_mode && it < max_iter) {
        ++it;
        // f!(x1, x0): H_MF[rho0]; mu; rho1 and eps_band  (selfconsistent.jl:281-295, :374-381)
        if ((st = hf_apply(S, S.rho0))) return st;
        // one factorization for mu and rho when the grid is one chunk (FP64 eigenvalues);
        // else bands, mu, then rho in a second pass.  For T > 0 on the per-k kernels mu stays
        // on the device (scal[8], scal[9]) and comes back with the iteration's scalars.
        const MuReq mq{filling, &mu, S.scal + 8};
        st = run_rho(S.hmf, Nk, ksd, kwd, 0.0, T, nRr, Rr, reinterpret_cast<double*>(S.rho1), S.scal + 4,
                     S.bands, nullptr, &mq);
        if (st == kFallback) {
            if ((st = run_bands(S.hmf, Nk, ksd, S.bands, nullptr))) return st;
            if ((st = mu_from_bands(c, N, Nk, S.bands, kwd, filling, T, &mu))) return st;
            st = run_rho(S.hmf, Nk, ksd, kwd, mu, T, nRr, Rr, reinterpret_cast<double*>(S.rho1), S.scal + 4,
                         nullptr);
        }
        if (st) return st;
        {
            StageScope scm(c, ST_MIX);
            hipLaunchKernelGGL(lqm::k_mix, dim3(kMixBlocks), dim3(256), 0, c->stream, nrho, beta, S.rho1,
                               S.rho0, S.part);
            hipLaunchKernelGGL(lqm::k_fold2, dim3(1), dim3(256), 0, c->stream, kMixBlocks, S.part,
                               S.scal + 5);
        }
        LQM_HIP(hipGetLastError());
        if (mq.on_device) {
            // one synchronisation for the scalars, mu and the kernels' status word
            LQM_HIP(hipMemcpyAsync(sc, S.scal, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
            LQM_HIP(hipMemcpyAsync(&hstat, c->d_status, sizeof(int), hipMemcpyDeviceToHost, c->stream));
            LQM_HIP(hipStreamSynchronize(c->stream));
            if (hstat) {
                if ((st = check_status(c, "densitymatrix"))) return st;
            }
            if (sc[9] == 2.0) return fail(LQM_ERR_NONFINITE, "non-finite bands");
            if (sc[9] == 1.0) return fail(LQM_ERR_ARG, "filling not bracketed by the band range");
            mu = sc[8];
        } else if ((st = d2h(c, sc, S.scal, sizeof(sc)))) {
            return st;
        }
        if ((st = hf_check_energy(S, sc))) return st;
        eband = sc[4];
        const double E = eband - (sc[0] + sc[1]);
        const double dn = std::sqrt(sc[5]), xn = std::sqrt(sc[6]);
        residual = xn > 2.220446049250313e-16 ? dn / xn : dn;
        if (history) {
            history[3 * (it - 1)] = E;
            history[3 * (it - 1) + 1] = residual;
            history[3 * (it - 1) + 2] = mu;
        }
        if (!std::isfinite(residual)) return fail(LQM_ERR_NONFINITE, "non-finite SCF residual");
        if (residual < tol) {
            conv = true;
            break;
        }
    }
    // refresh H_MF and the energies at the final rho (selfconsistent.jl:319-328)
    if ((st = hf_apply(S, S.rho0))) return st;
    if ((st = d2h(c, sc, S.scal, sizeof(sc)))) return st;
    if ((st = hf_check_energy(S, sc))) return st;
    const double dc = sc[0] + sc[1];
    result[0] = eband - dc;
    result[1] = mu;
    result[2] = eband;
    result[3] = sc[0];
    result[4] = sc[1];
    result[5] = eband - 2.0 * dc;
    result[6] = residual;
    result[7] = (double)it;
    *converged = conv ? 1 : 0;
    if ((st = d2h(c, rho, S.rho0, sizeof(cx<double>) * nrho))) return st;
    return hf_outputs(S, nK_out, RK_out, hmf_out);
}

lqm_status lqm_hartreefock_grandpotential(lqm_context* ctx, int N, int d, int nRh, const int32_t* Rh,
                                          const double* Th, int nRv, const int32_t* Rv, const double* Tv,
                                          int nRr, const int32_t* Rr, const double* rho, int flags, int Nk,
                                          const double* ks, const double* kweights, double mu, double T,
                                          double* omega, double* grad_h, double* grad_v) {
    if (Nk <= 0 || !ks) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (!(T >= 0.0)) return fail(LQM_ERR_ARG, "negative temperature");
    if (!omega) return fail(LQM_ERR_ARG, "null omega");
    HFState S;
    lqm_status st = hf_setup(S, ctx, N, d, nRh, Rh, Th, nRv, Rv, Tv, nRr, Rr, rho, flags, 0);
    if (st) return st;
    lqm_context* c = ctx;
    if ((st = hf_apply(S, S.rho0))) return st;
    void *dks, *dkw = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    const size_t NN = (size_t)N * N;
    const bool want = grad_h || grad_v;
    cx<double>* G = nullptr;
    if (want && (st = wsT(c, W_HFG, NN * S.nK, &G))) return st;
    if ((st = lqm_grandpotential_dev(S.hmf, Nk, (const double*)dks, (const double*)dkw, mu, T, omega,
                                     reinterpret_cast<double*>(G))))
        return st;
    if (!want) return LQM_OK;
    // key maps: h key r and v key l -> their mean-field key
    std::vector<int> maps((size_t)nRh + 2 * std::max(nRv, 1), 0);
    for (int q = 0; q < S.nK; ++q) {
        if (S.map[3 * q] >= 0) maps[S.map[3 * q]] = q;
        if (S.map[3 * q + 1] >= 0) maps[nRh + S.map[3 * q + 1]] = q;
    }
    for (int l = 0; l < nRv; ++l) maps[(size_t)nRh + std::max(nRv, 1) + l] = S.vr[l];
    const size_t nout = NN * ((size_t)nRh + std::max(nRv, 0));
    cx<double>* out;
    if ((st = wsT(c, W_HFA, nout + (maps.size() * sizeof(int) + 15) / 16 + 1, &out))) return st;
    int* dmaps = reinterpret_cast<int*>(out + nout);
    LQM_HIP(hipMemcpyAsync(dmaps, maps.data(), sizeof(int) * maps.size(), hipMemcpyHostToDevice, c->stream));
    {
        StageScope sc(c, ST_HF);
        hipLaunchKernelGGL(lqm::k_hf_adjoint, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, c->stream, N, nRh,
                           std::max(nRv, 0), dmaps, dmaps + nRh, dmaps + nRh + std::max(nRv, 1), S.qz, S.rz, S.flags,
                           G, S.rho0, out, out + NN * nRh);
    }
    LQM_HIP(hipGetLastError());
    if (grad_h && (st = d2h(c, grad_h, out, sizeof(cx<double>) * NN * nRh))) return st;
    if (grad_v && nRv > 0 && (st = d2h(c, grad_v, out + NN * nRh, sizeof(cx<double>) * NN * nRv))) return st;
    return LQM_OK;
}

lqm_status lqm_densitymatrix_occ_dev(lqm_model* m, int Nk, const double* ks_dev,
                                     const double* kweights_dev, const double* occ_dev, int nRr,
                                     const int32_t* Rr, double* rho_dev, double* bands_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (!occ_dev) return fail(LQM_ERR_ARG, "null occupations");
    double* eb;
    if ((st = wsT(m->ctx, W_SMALL, 8, &eb))) return st;
    return run_rho(m, Nk, ks_dev, kweights_dev, 0.0, 0.0, nRr, Rr, rho_dev, eb, bands_dev, occ_dev);
}

lqm_status lqm_densitymatrix_occ(lqm_model* m, int Nk, const double* ks, const double* kweights,
                                 const double* occ, int nRr, const int32_t* Rr, double* rho) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (!occ || (!rho && nRr > 0)) return fail(LQM_ERR_ARG, "null occupations or rho");
    lqm_context* c = m->ctx;
    const size_t NN = (size_t)m->N * m->N;
    void *dks, *dkw = nullptr, *docc, *drho;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = stage_in(c, W_OM, occ, sizeof(double) * m->N * Nk, &docc))) return st;
    if ((st = ws(c, W_OUT, sizeof(cx<double>) * NN * std::max(nRr, 1), &drho))) return st;
    if ((st = lqm_densitymatrix_occ_dev(m, Nk, (const double*)dks, (const double*)dkw,
                                        (const double*)docc, nRr, Rr, (double*)drho, nullptr)))
        return st;
    if (nRr > 0) return d2h(c, rho, drho, sizeof(cx<double>) * NN * nRr);
    return LQM_OK;
}

lqm_status lqm_grandpotential_dev(lqm_model* m, int Nk, const double* ks_dev,
                                  const double* kweights_dev, double mu, double T, double* omega,
                                  double* grad_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks_dev) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (!(T >= 0.0)) return fail(LQM_ERR_ARG, "negative temperature");
    if (!omega) return fail(LQM_ERR_ARG, "null omega");
    lqm_context* c = m->ctx;
    const int N = m->N, nR = m->nR;
    const size_t NN = (size_t)N * N;
    // the adjoint is rho on the reflected keys, transposed block by block
    std::vector<int32_t> negR(m->R.size());
    for (size_t i = 0; i < negR.size(); ++i) negR[i] = -m->R[i];
    cx<double>* tmp;
    double *bands, *small;
    if ((st = wsT(c, W_GRADT, grad_dev ? NN * nR : 1, &tmp))) return st;
    if ((st = wsT(c, W_BANDS, (size_t)N * Nk, &bands))) return st;
    const int Ne = N * Nk;
    const int G = std::max(1, std::min(1024, (Ne + 255) / 256));
    if ((st = wsT(c, W_SMALL, (size_t)G + 8, &small))) return st;
    if (grad_dev) {
        if ((st = run_rho(m, Nk, ks_dev, kweights_dev, mu, T, nR, negR.data(), reinterpret_cast<double*>(tmp),
                          small + G + 1, bands)))
            return st;
    } else if ((st = run_bands(m, Nk, ks_dev, bands, nullptr))) {
        return st;
    }
    {
        StageScope sc(c, ST_REDUCE);
        hipLaunchKernelGGL(lqm::k_omega, dim3(G), dim3(256), 0, c->stream, N, Ne, bands, kweights_dev, mu,
                           T, small);
        hipLaunchKernelGGL(lqm::k_sum, dim3(1), dim3(256), 0, c->stream, G, small, small + G, 0);
        if (grad_dev) {
            const size_t tot = NN * nR;
            hipLaunchKernelGGL(lqm::k_block_transpose, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                               c->stream, N, nR, tmp, reinterpret_cast<cx<double>*>(grad_dev));
        }
    }
    LQM_HIP(hipGetLastError());
    double h = 0.0;
    if ((st = d2h(c, &h, small + G, sizeof(double)))) return st;
    *omega = kweights_dev ? h : h / (double)Nk;
    return LQM_OK;
}

lqm_status lqm_grandpotential(lqm_model* m, int Nk, const double* ks, const double* kweights,
                              double mu, double T, double* omega, double* grad) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks) return fail(LQM_ERR_ARG, "Nk must be positive");
    lqm_context* c = m->ctx;
    const size_t NN = (size_t)m->N * m->N;
    void *dks, *dkw = nullptr, *dg = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if (grad && (st = ws(c, W_OUT, sizeof(cx<double>) * NN * m->nR, &dg))) return st;
    if ((st = lqm_grandpotential_dev(m, Nk, (const double*)dks, (const double*)dkw, mu, T, omega,
                                     (double*)dg)))
        return st;
    if (grad) return d2h(c, grad, dg, sizeof(cx<double>) * NN * m->nR);
    return LQM_OK;
}

lqm_status lqm_bandmatrix_expval_dev(lqm_model* m, int Nk, const double* ks_dev, int nP,
                                     lqm_model* const* projectors, double* bands_dev,
                                     double* obs_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk < 0 || nP < 0) return fail(LQM_ERR_ARG, "Nk and nP must be non-negative");
    if (Nk == 0) return LQM_OK;
    if (!ks_dev || !bands_dev || (nP > 0 && (!projectors || !obs_dev)))
        return fail(LQM_ERR_ARG, "null buffer");
    const int N = m->N;
    for (int p = 0; p < nP; ++p) {
        if (!projectors[p] || projectors[p]->ctx != m->ctx) return fail(LQM_ERR_ARG, "bad projector handle");
        if (projectors[p]->N != N || projectors[p]->d != m->d)
            return fail(LQM_ERR_ARG, "projector and Hamiltonian differ in N or d");
    }
    if (nP == 0) return run_bands(m, Nk, ks_dev, bands_dev, nullptr);
    lqm_context* c = m->ctx;
    const size_t NN = (size_t)N * N;
    const size_t budget = size_t(4) << 30;
    const int Kx = (int)std::max<size_t>(1, std::min<size_t>((size_t)Nk, budget / (3 * NN * sizeof(cx<double>))));
    cx<double> *V, *O, *W;
    if ((st = wsT(c, W_PV, NN * Kx, &V))) return st;
    if ((st = wsT(c, W_PO, NN * Kx, &O))) return st;
    if ((st = wsT(c, W_PW, NN * Kx, &W))) return st;
    const cx<double> one = {1.0, 0.0}, zero = {0.0, 0.0};
    for (int k0 = 0; k0 < Nk; k0 += Kx) {
        const int Kc = std::min(Kx, Nk - k0);
        const double* ksc = ks_dev + (size_t)k0 * m->d;
        if ((st = run_bands(m, Kc, ksc, bands_dev + (size_t)k0 * N, reinterpret_cast<double*>(V)))) return st;
        for (int p = 0; p < nP; ++p) {
            lqm_model* q = projectors[p];
            const size_t sh = sizeof(cx<double>) * q->nR;
            auto kH = lqm::k_fouriersum;
            allow_lds(kH, sh);
            hipLaunchKernelGGL(kH, dim3(Kc), dim3(256), sh, c->stream, N, q->d, q->nR, q->dR, q->dT, ksc, O);
            zgemm<0, 0>(c, N, Kc, O, V, W, one, zero);
            hipLaunchKernelGGL(lqm::k_expval_cols, dim3(Kc), dim3(256), 0, c->stream, N, V, W,
                               obs_dev + ((size_t)p * Nk + k0) * N);
            LQM_HIP(hipGetLastError());
        }
    }
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_bandmatrix_expval(lqm_model* m, int Nk, const double* ks, int nP,
                                 lqm_model* const* projectors, double* bands, double* obs) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk < 0 || nP < 0) return fail(LQM_ERR_ARG, "Nk and nP must be non-negative");
    if (Nk == 0) return LQM_OK;
    if (!ks || !bands || (nP > 0 && !obs)) return fail(LQM_ERR_ARG, "null buffer");
    lqm_context* c = m->ctx;
    const size_t N = m->N;
    void *dks, *db, *dobs = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if ((st = ws(c, W_BANDS, sizeof(double) * N * Nk, &db))) return st;
    if (nP > 0 && (st = ws(c, W_OUT, sizeof(double) * N * Nk * nP, &dobs))) return st;
    if ((st = lqm_bandmatrix_expval_dev(m, Nk, (const double*)dks, nP, projectors, (double*)db,
                                        (double*)dobs)))
        return st;
    if ((st = d2h(c, bands, db, sizeof(double) * N * Nk))) return st;
    if (nP > 0) return d2h(c, obs, dobs, sizeof(double) * N * Nk * nP);
    return LQM_OK;
}

// ---- consumers of the eigenvectors: ldos, density, Kubo ----------------------------
// Each pass walks the grid in chunks: the band pass leaves eigenvalues and eigenvectors of the
// chunk in HBM, then the consumer kernels read them; nbuf = N x N complex buffers per k-point.
static int consumer_chunk(int N, int Nk, int nbuf) {
    const size_t per = (size_t)nbuf * N * N * sizeof(cx<double>) + 16 * (size_t)N;
    return (int)std::max<size_t>(1, std::min<size_t>((size_t)Nk, (size_t(4) << 30) / per));
}

static lqm_status site_weights_dev(lqm_model* m, int Nk, const double* ks_dev, const double* kw_dev, int kind, int Nw,
                            const double* om_dev, double gamma, double mu, double* out_dev) {
    lqm_context* c = m->ctx;
    const int N = m->N;
    const size_t NN = (size_t)N * N;
    const int Kx = consumer_chunk(N, Nk, 1);
    const int G = std::min(Kx, 1024);
    cx<double>* V;
    double *eb, *part;
    lqm_status st;
    if ((st = wsT(c, W_PV, NN * Kx, &V))) return st;
    if ((st = wsT(c, W_RB, (size_t)N * Kx, &eb))) return st;
    if ((st = wsT(c, W_RP, (size_t)N * G, &part))) return st;
    const size_t sh = sizeof(double) * N;
    auto kS = lqm::k_site_weights;
    allow_lds(kS, sh);
    for (int k0 = 0; k0 < Nk; k0 += Kx) {
        const int Kc = std::min(Kx, Nk - k0);
        if ((st = run_bands(m, Kc, ks_dev + (size_t)k0 * m->d, eb, reinterpret_cast<double*>(V)))) return st;
        StageScope sc(c, ST_RESP, Kc);
        hipLaunchKernelGGL(kS, dim3(G), dim3(256), sh, c->stream, N, Kc, eb, V, kw_dev ? kw_dev + k0 : nullptr,
                           kind, Nw, om_dev, gamma, mu, k0 == 0 ? 1 : 0, part);
        LQM_HIP(hipGetLastError());
    }
    double scale = kw_dev ? 1.0 : 1.0 / (double)Nk;
    if (kind == 0) scale /= M_PI;
    {
        StageScope sc(c, ST_REDUCE);
        hipLaunchKernelGGL(lqm::k_dos_fold, dim3((N + 255) / 256), dim3(256), 0, c->stream, N, G, part, scale,
                           out_dev);
    }
    LQM_HIP(hipGetLastError());
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_ldos_dev(lqm_model* m, int Nk, const double* ks_dev, const double* kweights_dev, int Nw,
                        const double* omegas_dev, double gamma, double* ldos_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks_dev) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (Nw <= 0 || !omegas_dev) return fail(LQM_ERR_ARG, "at least one frequency");
    if (!(gamma > 0.0)) return fail(LQM_ERR_ARG, "the broadening must be positive");
    if (!ldos_dev) return fail(LQM_ERR_ARG, "null output");
    return site_weights_dev(m, Nk, ks_dev, kweights_dev, 0, Nw, omegas_dev, gamma, 0.0, ldos_dev);
}

lqm_status lqm_ldos(lqm_model* m, int Nk, const double* ks, const double* kweights, int Nw, const double* omegas,
                    double gamma, double* ldos) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks || Nw <= 0 || !omegas || !ldos) return fail(LQM_ERR_ARG, "bad sizes or null buffer");
    lqm_context* c = m->ctx;
    void *dks, *dkw = nullptr, *dom, *dout;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = stage_in(c, W_OM, omegas, sizeof(double) * Nw, &dom))) return st;
    if ((st = ws(c, W_OUT, sizeof(double) * m->N, &dout))) return st;
    if ((st = lqm_ldos_dev(m, Nk, (const double*)dks, (const double*)dkw, Nw, (const double*)dom, gamma,
                           (double*)dout)))
        return st;
    return d2h(c, ldos, dout, sizeof(double) * m->N);
}

lqm_status lqm_density_dev(lqm_model* m, int Nk, const double* ks_dev, const double* kweights_dev, double mu,
                           double* n_dev) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks_dev) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (!n_dev) return fail(LQM_ERR_ARG, "null output");
    return site_weights_dev(m, Nk, ks_dev, kweights_dev, 1, 0, nullptr, 0.0, mu, n_dev);
}

lqm_status lqm_density(lqm_model* m, int Nk, const double* ks, const double* kweights, double mu, double* n) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks || !n) return fail(LQM_ERR_ARG, "bad sizes or null buffer");
    lqm_context* c = m->ctx;
    void *dks, *dkw = nullptr, *dout;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = ws(c, W_OUT, sizeof(double) * m->N, &dout))) return st;
    if ((st = lqm_density_dev(m, Nk, (const double*)dks, (const double*)dkw, mu, (double*)dout))) return st;
    return d2h(c, n, dout, sizeof(double) * m->N);
}

lqm_status lqm_kubo_dev(lqm_model* h, lqm_model* j1, lqm_model* j2, int Nk, const double* ks_dev,
                        const double* kweights_dev, int Nw, const double* omegas_dev, double mu, double gamma,
                        double T, double* chi_dev) {
    lqm_status st = check_model(h);
    if (st) return st;
    if (Nk <= 0 || !ks_dev) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (Nw <= 0 || !omegas_dev || !chi_dev) return fail(LQM_ERR_ARG, "bad frequencies or null output");
    if (!(T >= 0.0)) return fail(LQM_ERR_ARG, "negative temperature");
    for (lqm_model* q : {j1, j2}) {
        if (!q || q->ctx != h->ctx) return fail(LQM_ERR_ARG, "bad current-operator handle");
        if (q->N != h->N || q->d != h->d) return fail(LQM_ERR_ARG, "current operator and Hamiltonian differ in N or d");
    }
    lqm_context* c = h->ctx;
    const int N = h->N;
    const size_t NN = (size_t)N * N;
    const bool same = j1 == j2;
    const int Kx = consumer_chunk(N, Nk, same ? 4 : 5);
    const int G = std::min(Kx, 1024);
    cx<double> *V, *O, *W, *A, *C, *part;
    double *eb, *occ;
    if ((st = wsT(c, W_PV, NN * Kx, &V))) return st;
    if ((st = wsT(c, W_PO, NN * Kx, &O))) return st;
    if ((st = wsT(c, W_PW, NN * Kx, &W))) return st;
    if ((st = wsT(c, W_RA, NN * Kx, &A))) return st;
    if (same) C = A;
    else if ((st = wsT(c, W_RC, NN * Kx, &C))) return st;
    if ((st = wsT(c, W_RB, (size_t)N * Kx, &eb))) return st;
    if ((st = wsT(c, W_ROCC, (size_t)N * Kx, &occ))) return st;
    if ((st = wsT(c, W_RP, (size_t)G * Nw, &part))) return st;
    const cx<double> one = {1.0, 0.0}, zero = {0.0, 0.0};
    // frequencies per lane: 4 for wide windows (one LDS read of a pair serves 4 terms)
    const int F = Nw > 128 ? 4 : (Nw > 64 ? 2 : 1);
    const dim3 gk(G, (unsigned)((Nw + 64 * F - 1) / (64 * F)));
    auto kK = F == 4 ? lqm::k_kubo<4> : (F == 2 ? lqm::k_kubo<2> : lqm::k_kubo<1>);
    for (int k0 = 0; k0 < Nk; k0 += Kx) {
        const int Kc = std::min(Kx, Nk - k0);
        const double* ksc = ks_dev + (size_t)k0 * h->d;
        if ((st = run_bands(h, Kc, ksc, eb, reinterpret_cast<double*>(V)))) return st;
        // U^H J U for each current operator: J(k) assembled like H(k), W = J U on the matrix
        // cores, then U^H W
        for (int q = 0; q < (same ? 1 : 2); ++q) {
            lqm_model* J = q == 0 ? j1 : j2;
            // J(k) for the whole chunk as one product on the matrix cores: the stack of blocks
            // [NN x nR] times the phase table [nR x Kc]
            cx<double>* phj;
            if ((st = wsT(c, W_SPPH, (size_t)J->nR * Kc, &phj))) return st;
            {
                StageScope sc(c, ST_HK64);
                hipLaunchKernelGGL(lqm::k_sp_phases, dim3((Kc * J->nR + 255) / 256), dim3(256), 0, c->stream, Kc,
                                   J->d, J->nR, J->dR, ksc, phj);
            }
            zgemm_gen<0, 0>(c, ST_HK64, Kc, 1,
                            gemm_args((int)NN, Kc, J->nR, J->dT, (int)NN, 0, phj, J->nR, 0, O, (int)NN, 0, one, zero));
            zgemm<0, 0>(c, N, Kc, O, V, W, one, zero);
            zgemm_gen<1, 0>(c, ST_GEMM_CN, Kc, Kc,
                            gemm_args(N, N, N, V, N, (long long)NN, W, N, (long long)NN, q == 0 ? A : C, N,
                                      (long long)NN, one, zero));
            LQM_HIP(hipGetLastError());
        }
        StageScope sc(c, ST_RESP, Kc);
        const int L = N * Kc;
        hipLaunchKernelGGL(lqm::k_fd_occ, dim3((L + 255) / 256), dim3(256), 0, c->stream, L, eb, mu, T, occ);
        hipLaunchKernelGGL(kK, gk, dim3(256), 0, c->stream, N, Kc, eb, occ, A, C,
                           kweights_dev ? kweights_dev + k0 : nullptr, Nw, omegas_dev, gamma, k0 == 0 ? 1 : 0,
                           part);
        LQM_HIP(hipGetLastError());
    }
    {
        StageScope sc(c, ST_REDUCE);
        hipLaunchKernelGGL(lqm::k_kubo_fold, dim3((Nw + 255) / 256), dim3(256), 0, c->stream, Nw, G, part,
                           kweights_dev ? 1.0 : 1.0 / (double)Nk, reinterpret_cast<cx<double>*>(chi_dev));
    }
    LQM_HIP(hipGetLastError());
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_kubo(lqm_model* h, lqm_model* j1, lqm_model* j2, int Nk, const double* ks, const double* kweights,
                    int Nw, const double* omegas, double mu, double gamma, double T, double* chi) {
    lqm_status st = check_model(h);
    if (st) return st;
    if (Nk <= 0 || !ks || Nw <= 0 || !omegas || !chi) return fail(LQM_ERR_ARG, "bad sizes or null buffer");
    lqm_context* c = h->ctx;
    void *dks, *dkw = nullptr, *dom, *dout;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * h->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = stage_in(c, W_OM, omegas, sizeof(double) * Nw, &dom))) return st;
    if ((st = ws(c, W_OUT, sizeof(cx<double>) * Nw, &dout))) return st;
    if ((st = lqm_kubo_dev(h, j1, j2, Nk, (const double*)dks, (const double*)dkw, Nw, (const double*)dom, mu,
                           gamma, T, (double*)dout)))
        return st;
    return d2h(c, chi, dout, sizeof(cx<double>) * Nw);
}

lqm_status lqm_bdg_filling(lqm_model* m, int Nk, const double* ks, const double* kweights, double mu,
                           double T, double* filling) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks || !filling) return fail(LQM_ERR_ARG, "bad k-points or null output");
    if (!(T >= 0.0)) return fail(LQM_ERR_ARG, "negative temperature");
    lqm_context* c = m->ctx;
    void *dks, *dkw = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    BdgScope S;
    if ((st = bdg_begin(m, S))) return st;
    return bdg_filling(S, Nk, (const double*)dks, (const double*)dkw, mu, T, filling);
}

lqm_status lqm_bdg_chemicalpotential(lqm_model* m, int Nk, const double* ks, const double* kweights,
                                     double target, double T, double lo, double hi, double* mu,
                                     int* passes) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks || !mu) return fail(LQM_ERR_ARG, "bad k-points or null output");
    if (!(T >= 0.0)) return fail(LQM_ERR_ARG, "negative temperature");
    if (!(lo < hi) || !std::isfinite(lo) || !std::isfinite(hi)) return fail(LQM_ERR_ARG, "bad bracket");
    lqm_context* c = m->ctx;
    void *dks, *dkw = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    BdgScope S;
    if ((st = bdg_begin(m, S))) return st;
    int np = 0;
    auto f = [&](double x, double* out) {
        double fl = 0.0;
        lqm_status s2 = bdg_filling(S, Nk, (const double*)dks, (const double*)dkw, x, T, &fl);
        ++np;
        *out = fl - target;
        if (!s2 && !std::isfinite(*out)) return fail(LQM_ERR_NONFINITE, "non-finite BdG filling");
        return s2;
    };
    double a = lo, b = hi, fa, fb, fm;
    if ((st = f(a, &fa)) || (st = f(b, &fb))) return st;
    if (passes) *passes = np;
    if (fa == 0.0) return (*mu = a, LQM_OK);
    if (fb == 0.0) return (*mu = b, LQM_OK);
    if ((fa < 0.0) == (fb < 0.0)) return fail(LQM_ERR_ARG, "the target filling is not bracketed by [lo, hi]");
    for (int it = 0; it < 200; ++it) {
        const double x = roots_middle(a, b);
        if (!(a < x && x < b)) break;
        if ((st = f(x, &fm))) return st;
        if (fm == 0.0) {
            a = b = x;
            fa = fb = 0.0;
            break;
        }
        if ((fm < 0.0) == (fa < 0.0)) {
            a = x;
            fa = fm;
        } else {
            b = x;
            fb = fm;
        }
    }
    *mu = std::fabs(fa) <= std::fabs(fb) ? a : b;
    if (passes) *passes = np;
    return LQM_OK;
}

lqm_status lqm_bdg_densitymatrix(lqm_model* m, int Nk, const double* ks, const double* kweights, double mu,
                                 double T, int nRr, const int32_t* Rr, double* rho, double* ekin) {
    lqm_status st = check_model(m);
    if (st) return st;
    if (Nk <= 0 || !ks) return fail(LQM_ERR_ARG, "Nk must be positive");
    if (nRr < 0 || (nRr > 0 && (!Rr || !rho))) return fail(LQM_ERR_ARG, "bad rho keys or null rho");
    lqm_context* c = m->ctx;
    const size_t NN = (size_t)m->N * m->N;
    const int d = m->N / 2;
    void *dks, *dkw = nullptr, *drho;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if (kweights && (st = stage_in(c, W_KW, kweights, sizeof(double) * Nk, &dkw))) return st;
    if ((st = ws(c, W_OUT, sizeof(cx<double>) * NN * std::max(nRr, 1) + 16, &drho))) return st;
    double* deb = reinterpret_cast<double*>(static_cast<char*>(drho) + sizeof(cx<double>) * NN * std::max(nRr, 1));
    double eb = 0.0, trh = 0.0;
    {
        BdgScope S;
        if ((st = bdg_begin(m, S))) return st;
        // Re tr of the electron block of the unshifted zero key, for the Nambu constant
        std::vector<cx<double>> diag(d);
        LQM_HIP(hipMemcpy2DAsync(diag.data(), sizeof(cx<double>), S.base, sizeof(cx<double>) * (m->N + 1),
                                 sizeof(cx<double>), d, hipMemcpyDeviceToHost, c->stream));
        if ((st = bdg_shift(S, mu))) return st;
        if ((st = run_rho(m, Nk, (const double*)dks, (const double*)dkw, 0.0, T, nRr, Rr, (double*)drho, deb, nullptr)))
            return st;
        if ((st = d2h(c, &eb, deb, sizeof(double)))) return st;
        for (int i = 0; i < d; ++i) trh += diag[i].re;
    }
    if (nRr > 0 && (st = d2h(c, rho, drho, sizeof(cx<double>) * NN * nRr))) return st;
    if (ekin) *ekin = eb / 2.0 + (trh - d * mu) / 2.0;
    return LQM_OK;
}

lqm_status lqm_link_overlaps_dev(lqm_context* c, int N, int M, const double* U_dev, int nlink,
                                 const int32_t* pairs, double* O_dev) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (N <= 0 || M <= 0 || M > N || nlink < 0) return fail(LQM_ERR_ARG, "bad sizes");
    if (nlink == 0) return LQM_OK;
    if (nlink > 65535) return fail(LQM_ERR_UNSUPPORTED, "at most 65535 links per call");
    if (!U_dev || !pairs || !O_dev) return fail(LQM_ERR_ARG, "null buffer");
    for (int l = 0; l < 2 * nlink; ++l)
        if (pairs[l] < 0) return fail(LQM_ERR_ARG, "negative state index in the link list");
    int* dp;
    lqm_status st;
    if ((st = wsT(c, W_LNK, (size_t)2 * nlink, &dp))) return st;
    LQM_HIP(hipMemcpyAsync(dp, pairs, sizeof(int) * 2 * nlink, hipMemcpyHostToDevice, c->stream));
    const cx<double>* U = reinterpret_cast<const cx<double>*>(U_dev);
    lqm::ZGemmArgs g = gemm_args(M, M, N, U, N, (long long)N * M, U, N, (long long)N * M,
                                 reinterpret_cast<cx<double>*>(O_dev), M, (long long)M * M, {1.0, 0.0}, {0.0, 0.0});
    g.pairs = dp;
    zgemm_gen<1, 0>(c, ST_REDUCE, nlink, nlink, g);
    LQM_HIP(hipGetLastError());
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_link_phases_dev(lqm_context* c, int N, int M, const double* U_dev, int nlink, const int32_t* pairs,
                               double* phase_dev) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (nlink == 0) return LQM_OK;
    if (!phase_dev) return fail(LQM_ERR_ARG, "null buffer");
    if (M <= 0 || nlink < 0) return fail(LQM_ERR_ARG, "bad sizes");
    cx<double>* O;
    lqm_status st;
    if ((st = wsT(c, W_LNKO, (size_t)M * M * std::max(nlink, 1), &O))) return st;
    if ((st = lqm_link_overlaps_dev(c, N, M, U_dev, nlink, pairs, reinterpret_cast<double*>(O)))) return st;
    hipLaunchKernelGGL(lqm::k_det_phase, dim3(nlink), dim3(256), 0, c->stream, M, O, phase_dev);
    LQM_HIP(hipGetLastError());
    LQM_HIP(hipStreamSynchronize(c->stream));
    return LQM_OK;
}

lqm_status lqm_link_phases(lqm_context* c, int N, int M, int npts, const double* U, int nlink,
                           const int32_t* pairs, double* phase) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (N <= 0 || M <= 0 || M > N || npts <= 0 || nlink < 0) return fail(LQM_ERR_ARG, "bad sizes");
    if (nlink == 0) return LQM_OK;
    if (!U || !pairs || !phase) return fail(LQM_ERR_ARG, "null buffer");
    for (int l = 0; l < 2 * nlink; ++l)
        if (pairs[l] < 0 || pairs[l] >= npts) return fail(LQM_ERR_ARG, "state index out of range in the link list");
    void *dU, *dph;
    lqm_status st;
    if ((st = stage_in(c, W_KS, U, sizeof(cx<double>) * (size_t)N * M * npts, &dU))) return st;
    if ((st = ws(c, W_OUT, sizeof(double) * 2 * nlink, &dph))) return st;
    if ((st = lqm_link_phases_dev(c, N, M, (const double*)dU, nlink, pairs, (double*)dph))) return st;
    return d2h(c, phase, dph, sizeof(double) * 2 * nlink);
}

lqm_status lqm_link_overlaps(lqm_context* c, int N, int M, int npts, const double* U, int nlink,
                             const int32_t* pairs, double* O) {
    if (!c) return fail(LQM_ERR_ARG, "null context");
    if (N <= 0 || M <= 0 || M > N || npts <= 0 || nlink < 0) return fail(LQM_ERR_ARG, "bad sizes");
    if (nlink == 0) return LQM_OK;
    if (!U || !pairs || !O) return fail(LQM_ERR_ARG, "null buffer");
    for (int l = 0; l < 2 * nlink; ++l)
        if (pairs[l] < 0 || pairs[l] >= npts) return fail(LQM_ERR_ARG, "state index out of range in the link list");
    void *dU, *dO;
    lqm_status st;
    if ((st = stage_in(c, W_KS, U, sizeof(cx<double>) * (size_t)N * M * npts, &dU))) return st;
    if ((st = ws(c, W_OUT, sizeof(cx<double>) * (size_t)M * M * nlink, &dO))) return st;
    if ((st = lqm_link_overlaps_dev(c, N, M, (const double*)dU, nlink, pairs, (double*)dO))) return st;
    return d2h(c, O, dO, sizeof(cx<double>) * (size_t)M * M * nlink);
}

// ---- partial spectrum: Chebyshev-filtered subspace iteration ------------------------
lqm_status lqm_sparse_model_create(lqm_context* ctx, int N, int d, int nR, const int32_t* R, int64_t nnz,
                                   const int64_t* rowptr, const int32_t* col, const int32_t* key, const double* val,
                                   lqm_sparse_model** out) {
    if (!ctx || !out) return fail(LQM_ERR_ARG, "null context or output pointer");
    if (N <= 0 || d <= 0 || nR <= 0 || nnz < 0) return fail(LQM_ERR_ARG, "N, d, nR must be positive, nnz >= 0");
    if (nR > lqm::kSpMaxKeys) return fail(LQM_ERR_UNSUPPORTED, "more than 2048 keys");
    if (!R || !rowptr || (nnz > 0 && (!col || !key || !val))) return fail(LQM_ERR_ARG, "null array");
    if (rowptr[0] != 0 || rowptr[N] != nnz) return fail(LQM_ERR_ARG, "row pointer must run from 0 to nnz");
    std::vector<int2> ck((size_t)std::max<int64_t>(nnz, 1));
    double bound = 0.0;
    for (int i = 0; i < N; ++i) {
        if (rowptr[i + 1] < rowptr[i]) return fail(LQM_ERR_ARG, "row pointer decreases");
        double rs = 0.0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            if (col[e] < 0 || col[e] >= N || key[e] < 0 || key[e] >= nR)
                return fail(LQM_ERR_ARG, "column or key index out of range");
            if (!std::isfinite(val[2 * e]) || !std::isfinite(val[2 * e + 1])) return fail(LQM_ERR_NONFINITE, "non-finite value");
            ck[e] = make_int2(col[e], key[e]);
            rs += std::hypot(val[2 * e], val[2 * e + 1]);
        }
        bound = std::max(bound, rs);
    }
    auto* m = new lqm_sparse_model();
    m->ctx = ctx;
    m->N = N;
    m->d = d;
    m->nR = nR;
    m->nnz = nnz;
    m->R.assign(R, R + (size_t)d * nR);
    m->bound = bound;
    const size_t ne = (size_t)std::max<int64_t>(nnz, 1);
    hipError_t e = dev_malloc(&m->dR, sizeof(int) * d * nR);
    if (e == hipSuccess) e = dev_malloc(&m->rowptr, sizeof(long long) * (N + 1));
    if (e == hipSuccess) e = dev_malloc(&m->colkey, sizeof(int2) * ne);
    if (e == hipSuccess) e = dev_malloc(&m->val, sizeof(cx<double>) * ne);
    if (e == hipSuccess) e = hipMemcpyAsync(m->dR, R, sizeof(int) * d * nR, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(m->rowptr, rowptr, sizeof(long long) * (N + 1), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && nnz > 0)
        e = hipMemcpyAsync(m->colkey, ck.data(), sizeof(int2) * nnz, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && nnz > 0)
        e = hipMemcpyAsync(m->val, val, sizeof(cx<double>) * nnz, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)dev_free(m->dR);
        (void)dev_free(m->rowptr);
        (void)dev_free(m->colkey);
        (void)dev_free(m->val);
        delete m;
        return fail(LQM_ERR_DEVICE, std::string("sparse model allocation: ") + hipGetErrorString(e));
    }
    *out = m;
    return LQM_OK;
}

lqm_status lqm_sparse_model_destroy(lqm_sparse_model* m) {
    if (!m) return LQM_OK;
    (void)hipStreamSynchronize(m->ctx->stream);
    (void)dev_free(m->dR);
    (void)dev_free(m->rowptr);
    (void)dev_free(m->colkey);
    (void)dev_free(m->val);
    delete m;
    return LQM_OK;
}

lqm_status lqm_sparse_model_bound(lqm_sparse_model* m, double* bound) {
    if (!m || !bound) return fail(LQM_ERR_ARG, "null model or output");
    *bound = m->bound;
    return LQM_OK;
}

namespace {

// Eigenpairs of Kc Hermitian n x n matrices on the device (A, column-major, upper triangle
// read): the per-k eigensolver run on them through a stand-in model of one zero key (the
// variants that read a materialised H(k)).  theta: ascending eigenvalues, P: eigenvectors.
lqm_status eigh_given(lqm_model* dm, int Kc, const double* ks0, const cx<double>* A, double* theta, cx<double>* P) {
    lqm_context* c = dm->ctx;
    const int n = dm->N;
    Plan PL = make_plan(c, n, 1, Kc, true, false);
    PL.nqA = 0;  // the one-wave tridiagonalizations assemble H(k) themselves
    lqm_status st;
    for (int k0 = 0; k0 < Kc; k0 += PL.Kc) {
        const int kc = std::min(PL.Kc, Kc - k0);
        const size_t o = (size_t)k0 * n * n;
        if ((st = launch_eigvals<double>(dm, PL, kc, ks0, theta + (size_t)k0 * n, true, A + o))) return st;
        if ((st = launch_backtransform<double, 0>(dm, PL, kc, theta + (size_t)k0 * n, nullptr, 0.0, 0.0, 0.0, P + o,
                                                  nullptr)))
            return st;
    }
    return LQM_OK;
}

struct ModelGuard {
    lqm_model* m = nullptr;
    ~ModelGuard() { lqm_model_destroy(m); }
};

}  // namespace

lqm_status lqm_eigen_partial_dev(lqm_sparse_model* m, int Nk, const double* ks_dev, int nb, int which, double sigma,
                                 double tol, int maxiter, double* evals_dev, double* evecs_dev, int32_t* iters) {
    if (!m) return fail(LQM_ERR_ARG, "null model");
    if (Nk <= 0 || !ks_dev || !evals_dev) return fail(LQM_ERR_ARG, "bad k-points or null output");
    const int N = m->N;
    if (nb < 1 || nb > N) return fail(LQM_ERR_ARG, "num_bands must be in 1..N");
    if (which < 0 || which > 2) return fail(LQM_ERR_ARG, "which: 0 closest to sigma, 1 smallest, 2 largest");
    if (!(tol > 0.0) || maxiter < 1) return fail(LQM_ERR_ARG, "tol must be positive, maxiter >= 1");
    if (!std::isfinite(sigma)) return fail(LQM_ERR_ARG, "non-finite sigma");
    const int msub = std::min(N, std::max(2 * nb + 10, 20) > 64 ? std::max(nb + 8, 64) : std::max(2 * nb + 10, 20));
    if (msub > 64) return fail(LQM_ERR_UNSUPPORTED, "num_bands above 56 (subspace wider than 64)");
    const int mp = msub <= 16 ? 16 : (msub <= 32 ? 32 : 64);
    lqm_context* c = m->ctx;
    begin_pass(c, LQM_PREC_FP64);
    const hipStream_t s = c->stream;
    const size_t NP = (size_t)N * mp, MM = (size_t)msub * msub;
    const size_t per = 4 * 16 * NP + 4 * 16 * MM + 64 * (size_t)msub + 16 * (size_t)m->nR + 64;
    const int Kx = (int)std::max<size_t>(1, std::min<size_t>({(size_t)Nk, (size_t(16) << 30) / per, (size_t)65535}));
    lqm_status st;
    cx<double> *buf, *ph, *mats;
    double *th, *coef, *ks0;
    int* ints;
    if ((st = wsT(c, W_SPB, 4 * NP * Kx, &buf))) return st;
    if ((st = wsT(c, W_SPPH, (size_t)m->nR * Kx, &ph))) return st;
    if ((st = wsT(c, W_SPM, 4 * MM * Kx, &mats))) return st;
    if ((st = wsT(c, W_SPT, 2 * (size_t)msub * Kx, &th))) return st;
    if ((st = wsT(c, W_SPI, 4 * (size_t)Kx, &ints))) return st;
    if ((st = wsT(c, W_SPKS, (size_t)Kx, &ks0))) return st;
    LQM_HIP(hipMemsetAsync(ks0, 0, sizeof(double) * Kx, s));
    cx<double>* G = mats;
    cx<double>* Li = mats + MM * Kx;
    cx<double>* S = mats + 2 * MM * Kx;
    cx<double>* P = mats + 3 * MM * Kx;
    double* rn = th + (size_t)msub * Kx;
    int* kidx = ints;
    int* flag = ints + Kx;
    int* lo_d = ints + 2 * Kx;
    int* outl = ints + 3 * Kx;
    // stand-in model for the Rayleigh-Ritz eigensolves
    ModelGuard dm;
    {
        const int32_t z[1] = {0};
        if ((st = model_init(c, msub, 1, 1, z, &dm.m))) return st;
        LQM_HIP(hipMemsetAsync(dm.m->dR, 0, sizeof(int), s));
        LQM_HIP(hipMemsetAsync(dm.m->dT, 0, sizeof(cx<double>) * MM, s));
    }
    // spectrum of B: LM (H - sigma)^2 in [0, beta^2]; SR H and LR -H in [-rho, rho]
    const double rho = std::max(m->bound, 1e-300);
    const double beta = rho + std::fabs(sigma);
    const double a0 = which == 0 ? 0.0 : -rho, ub = which == 0 ? beta * beta : rho;
    const double shift = which == 0 ? sigma : 0.0;
    const double acosh_kappa = std::acosh(1e7);
    const int nmax = 4000;
    const double shift_rel = 11.0 * ((double)N * msub + (double)msub * (msub + 1)) * std::ldexp(1.0, -53);
    const cx<double> one = {1.0, 0.0}, zero = {0.0, 0.0};
    const long long sB = (long long)NP, sM = (long long)MM;
    auto spmm = [&](int nact, const double* cf, const cx<double>* X, const cx<double>* pp, const cx<double>* qq,
                    cx<double>* Y) {
        hipLaunchKernelGGL(lqm::k_spmm, dim3((unsigned)((N + lqm::kSpRows - 1) / lqm::kSpRows), (unsigned)nact),
                           dim3(256), 0, s, N, mp, msub, m->rowptr, m->colkey, m->val, m->nR, ph, shift, kidx, cf, X,
                           pp, qq, Y);
    };
    std::vector<double> hth, hrn, hcoef;
    std::vector<int> hk, hlo;
    std::vector<double> cut;
    const size_t chol_lds = sizeof(cx<double>) * 65 * msub;
    auto kC = lqm::k_chol_inv;
    allow_lds(kC, chol_lds);
    for (int k0 = 0; k0 < Nk; k0 += Kx) {
        const int Kc = std::min(Kx, Nk - k0);
        const double* ksc = ks_dev + (size_t)k0 * m->d;
        hipLaunchKernelGGL(lqm::k_sp_phases, dim3((Kc * m->nR + 255) / 256), dim3(256), 0, s, Kc, m->d, m->nR, m->dR,
                           ksc, ph);
        cx<double>* B[4] = {buf, buf + NP * Kx, buf + 2 * NP * Kx, buf + 3 * NP * Kx};
        int ix = 0;  // the buffer holding the current block
        hipLaunchKernelGGL(lqm::k_sp_start, dim3((unsigned)std::min<size_t>(4096, (NP * Kc + 255) / 256)), dim3(256), 0,
                           s, N, mp, msub, Kc, (long long)k0, 0x1a771ce9ull, B[0]);
        LQM_HIP(hipGetLastError());
        hk.resize(Kc);
        for (int q = 0; q < Kc; ++q) hk[q] = q;
        cut.assign(Kc, 0.0);
        hlo.assign(Kc, 0);
        {
            // first cut: the B-value below which about 2m of the N eigenvalues lie for a flat
            // density of states, kept inside (a0, ub)
            const double fr = std::min(0.25, 2.0 * msub / (double)N);
            for (int q = 0; q < Kc; ++q)
                cut[q] = which == 0 ? std::max(1e-6, fr * fr) * ub : a0 + std::max(1e-6, 2.0 * fr) * (ub - a0);
        }
        std::vector<int> done_it(Kc, -1);
        int it = 0;
        for (; it < maxiter && !hk.empty(); ++it) {
            const int nact = (int)hk.size();
            LQM_HIP(hipMemcpyAsync(kidx, hk.data(), sizeof(int) * nact, hipMemcpyHostToDevice, s));
            // filter degree: the conditioning bound T_n(|tau0|) <= 1e7 of the most demanding slot
            int n = nmax;
            std::vector<double> ee(Kc), cc(Kc), t0(Kc);
            for (int q : hk) {
                const double a = std::min(std::max(cut[q], a0 + 1e-14 * (ub - a0)), ub - 1e-9 * (ub - a0));
                ee[q] = 0.5 * (ub - a);
                cc[q] = 0.5 * (ub + a);
                t0[q] = (a0 - cc[q]) / ee[q];
                const double ac = std::acosh(std::max(1.0 + 1e-300, std::fabs(t0[q])));
                n = std::min(n, std::max(1, (int)std::floor(acosh_kappa / ac)));
            }
            // per-step coefficients {alpha, beta1, beta2} of the scaled Chebyshev recurrence
            // q1 = (B - c) q0 / (e r0), q_{s+1} = 2 (B - c) q_s / (e r_s) - q_{s-1} / (r_s r_{s-1}),
            // r_0 = tau0, r_s = 2 tau0 - 1 / r_{s-1}: the component at a0 stays 1
            hcoef.assign((size_t)n * Kx * 3, 0.0);
            for (int q : hk) {
                double rprev = 0.0, r = t0[q];
                for (int j = 0; j < n; ++j) {
                    if (j > 0) {
                        rprev = r;
                        r = 2.0 * t0[q] - 1.0 / rprev;
                    }
                    const double al = (j == 0 ? 1.0 : 2.0) / (ee[q] * r);
                    double* cf = &hcoef[((size_t)j * Kx + q) * 3];
                    cf[0] = which == 2 ? -al : al;
                    cf[1] = -al * cc[q];
                    cf[2] = j == 0 ? 0.0 : -1.0 / (r * rprev);
                }
            }
            if ((st = wsT(c, W_SPC, hcoef.size(), &coef))) return st;
            LQM_HIP(hipMemcpyAsync(coef, hcoef.data(), sizeof(double) * hcoef.size(), hipMemcpyHostToDevice, s));
            {
                StageScope sc(c, ST_SPF, (long long)nact * n);
                // q_0 in B[ix]; q_1 goes to B[c1]; later steps overwrite q_{s-1} in place
                const int c1 = (ix + 1) % 4, tmp = (ix + 2) % 4;
                int prev = ix, cur = c1;
                for (int j = 0; j < n; ++j) {
                    const double* cf = coef + (size_t)j * Kx * 3;
                    const int src = j == 0 ? ix : cur;
                    const int dst = j == 0 ? c1 : prev;
                    const cx<double>* qq = j == 0 ? nullptr : B[prev];
                    if (which == 0) {
                        spmm(nact, nullptr, B[src], nullptr, nullptr, B[tmp]);
                        spmm(nact, cf, B[tmp], B[src], qq, B[dst]);
                    } else {
                        spmm(nact, cf, B[src], B[src], qq, B[dst]);
                    }
                    if (j > 0) {
                        prev = cur;
                        cur = dst;
                    }
                }
                ix = cur;
            }
            LQM_HIP(hipGetLastError());
            // orthonormal basis: shifted CholeskyQR, then two plain passes (Z <- L^-1 Z)
            for (int pass = 0; pass < 3; ++pass) {
                const int dst = (ix + 1) % 4;
                zgemm_gen<0, 1>(c, ST_GEMM_NC, nact, nact,
                                [&] {
                                    auto g = gemm_args(msub, msub, N, B[ix], mp, sB, B[ix], mp, sB, G, msub, sM, one, zero);
                                    g.kidx = kidx;
                                    return g;
                                }());
                hipLaunchKernelGGL(kC, dim3(nact), dim3(256), chol_lds, s, msub, kidx, G, shift_rel, pass == 0 ? 1 : 0,
                                   Li, flag);
                zgemm_gen<0, 0>(c, ST_GEMM_NN, nact, nact,
                                [&] {
                                    auto g = gemm_args(msub, N, msub, Li, msub, sM, B[ix], mp, sB, B[dst], mp, sB, one, zero);
                                    g.kidx = kidx;
                                    return g;
                                }());
                ix = dst;
            }
            LQM_HIP(hipGetLastError());
            // Rayleigh-Ritz: W = (H - shift) Y, conj(S) = Z W_z^H, its eigenpairs, then the Ritz
            // block X = P^H Z and W Q = P^H W_z, and the residual norms
            const int iw = (ix + 1) % 4, ixn = (ix + 2) % 4, iwq = (ix + 3) % 4;
            spmm(nact, nullptr, B[ix], nullptr, nullptr, B[iw]);
            zgemm_gen<0, 1>(c, ST_GEMM_NC, nact, nact,
                            [&] {
                                auto g = gemm_args(msub, msub, N, B[ix], mp, sB, B[iw], mp, sB, S, msub, sM, one, zero);
                                g.kidx = kidx;
                                return g;
                            }());
            LQM_HIP(hipGetLastError());
            if ((st = eigh_given(dm.m, Kc, ks0, S, th, P))) return st;
            zgemm_gen<1, 0>(c, ST_GEMM_CN, nact, nact,
                            [&] {
                                auto g = gemm_args(msub, N, msub, P, msub, sM, B[ix], mp, sB, B[ixn], mp, sB, one, zero);
                                g.kidx = kidx;
                                return g;
                            }());
            zgemm_gen<1, 0>(c, ST_GEMM_CN, nact, nact,
                            [&] {
                                auto g = gemm_args(msub, N, msub, P, msub, sM, B[iw], mp, sB, B[iwq], mp, sB, one, zero);
                                g.kidx = kidx;
                                return g;
                            }());
            hipLaunchKernelGGL(lqm::k_sp_resid, dim3(nact), dim3(256), 0, s, N, mp, msub, kidx, B[ixn], B[iwq], th, rn);
            LQM_HIP(hipGetLastError());
            ix = ixn;
            hth.resize((size_t)msub * Kc);
            hrn.resize((size_t)msub * Kc);
            LQM_HIP(hipMemcpyAsync(hth.data(), th, sizeof(double) * msub * Kc, hipMemcpyDeviceToHost, s));
            LQM_HIP(hipMemcpyAsync(hrn.data(), rn, sizeof(double) * msub * Kc, hipMemcpyDeviceToHost, s));
            LQM_HIP(hipStreamSynchronize(s));
            std::vector<int> keep, fin;
            for (int q : hk) {
                const double* t = &hth[(size_t)q * msub];
                int lo = 0;
                if (which == 2) {
                    lo = msub - nb;
                } else if (which == 0) {
                    // the nb Ritz values nearest 0 (a contiguous window of the ascending list)
                    int a = 0;
                    for (int i = 1; i < msub; ++i)
                        if (std::fabs(t[i]) < std::fabs(t[a])) a = i;
                    int l = a, r = a;
                    while (r - l + 1 < nb) {
                        if (l == 0) ++r;
                        else if (r == msub - 1) --l;
                        else if (std::fabs(t[l - 1]) <= std::fabs(t[r + 1])) --l;
                        else ++r;
                    }
                    lo = l;
                }
                hlo[q] = lo;
                bool conv = true;
                for (int b = lo; b < lo + nb; ++b) conv = conv && hrn[(size_t)q * msub + b] <= tol * rho;
                if (conv || it + 1 == maxiter) {
                    if (conv) done_it[q] = it + 1;
                    fin.push_back(q);
                    continue;
                }
                keep.push_back(q);
                double nc;
                if (which == 0) {
                    nc = 0.0;
                    for (int i = 0; i < msub; ++i) nc = std::max(nc, t[i] * t[i]);
                } else {
                    nc = which == 1 ? t[msub - 1] : -t[0];
                }
                cut[q] = nc;
            }
            // the finished slots (converged, or at the iteration limit) leave their eigenpairs now:
            // later iterations no longer update their block
            if (!fin.empty()) {
                LQM_HIP(hipMemcpyAsync(lo_d, hlo.data(), sizeof(int) * Kc, hipMemcpyHostToDevice, s));
                LQM_HIP(hipMemcpyAsync(outl, fin.data(), sizeof(int) * fin.size(), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL(lqm::k_sp_output,
                                   dim3((unsigned)std::min<size_t>(1024, ((size_t)N * nb + 255) / 256), (unsigned)fin.size()),
                                   dim3(256), 0, s, N, mp, msub, nb, outl, lo_d, th, shift, B[ix],
                                   evals_dev + (size_t)k0 * nb,
                                   evecs_dev ? reinterpret_cast<cx<double>*>(evecs_dev) + (size_t)k0 * N * nb : nullptr);
                LQM_HIP(hipGetLastError());
                LQM_HIP(hipStreamSynchronize(s));  // outl and lo_d are rewritten next iteration
            }
            hk.swap(keep);
        }
        if (iters)
            for (int q = 0; q < Kc; ++q) iters[k0 + q] = done_it[q];
    }
    LQM_HIP(hipStreamSynchronize(s));
    return check_status(c, "partial spectrum");
}

lqm_status lqm_eigen_partial(lqm_sparse_model* m, int Nk, const double* ks, int nb, int which, double sigma,
                             double tol, int maxiter, double* evals, double* evecs, int32_t* iters) {
    if (!m) return fail(LQM_ERR_ARG, "null model");
    if (Nk <= 0 || !ks || !evals) return fail(LQM_ERR_ARG, "bad k-points or null output");
    if (nb < 1 || nb > m->N) return fail(LQM_ERR_ARG, "num_bands must be in 1..N");
    lqm_context* c = m->ctx;
    lqm_status st;
    void *dks, *dev, *dvec = nullptr;
    if ((st = stage_in(c, W_KS, ks, sizeof(double) * m->d * Nk, &dks))) return st;
    if ((st = ws(c, W_OUT, sizeof(double) * (size_t)nb * Nk, &dev))) return st;
    if (evecs && (st = ws(c, W_OM, sizeof(cx<double>) * (size_t)m->N * nb * Nk, &dvec))) return st;
    if ((st = lqm_eigen_partial_dev(m, Nk, (const double*)dks, nb, which, sigma, tol, maxiter, (double*)dev,
                                    (double*)dvec, iters)))
        return st;
    if ((st = d2h(c, evals, dev, sizeof(double) * (size_t)nb * Nk))) return st;
    if (evecs) return d2h(c, evecs, dvec, sizeof(cx<double>) * (size_t)m->N * nb * Nk);
    return LQM_OK;
}

lqm_status lqm_hartreefock_scf_batch(lqm_context* ctx, int d, int M, lqm_hf_problem* problems,
                                     double T, double beta, double tol, int max_iter, int flags,
                                     int nstreams) {
    if (!ctx) return fail(LQM_ERR_ARG, "null context");
    if (M < 0 || (M > 0 && !problems)) return fail(LQM_ERR_ARG, "bad problem list");
    if (nstreams < 1 || nstreams > 16) return fail(LQM_ERR_ARG, "nstreams must be in 1..16");
    if (M == 0) return LQM_OK;
    // longest first (N^3 Nk), so the large models start early and small ones fill in
    std::vector<int> order(M);
    for (int i = 0; i < M; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        const double ca = std::pow((double)problems[a].N, 3) * problems[a].Nk;
        const double cb = std::pow((double)problems[b].N, 3) * problems[b].Nk;
        return ca > cb;
    });
    const int W = std::min(nstreams, M);
    // one HBM budget across the workers: each worker context's per-chunk scratch is capped at
    // an equal share of 85 % of the device memory free now (the largest models then run in
    // more chunks rather than together exceed the device)
    size_t cap = SIZE_MAX;
    {
        size_t fr = 0, tot = 0;
        LQM_HIP(hipSetDevice(ctx->device));
        LQM_HIP(hipMemGetInfo(&fr, &tot));
        cap = std::max<size_t>(size_t(256) << 20, (size_t)(0.85 * (double)fr) / (size_t)W);
    }
    std::atomic<int> next{0};
    std::mutex prog_mu;
    std::vector<std::string> errs(W);
    std::vector<int> errst(W, LQM_OK);
    auto worker = [&](int w) {
        lqm_context* wc = nullptr;
        lqm_status st = lqm_context_create(ctx->device, ctx->precision, &wc);
        if (st) {
            errst[w] = st;
            errs[w] = g_err;
            return;
        }
        wc->refine_max = ctx->refine_max;
        wc->refine_thr = ctx->refine_thr;
        wc->ws_cap = std::min(ctx->ws_cap, cap);
        wc->tuned = ctx->tuned;
        for (int k = 0; k < VK_N; ++k) wc->forced[k] = ctx->forced[k];
        for (int i = next.fetch_add(1); i < M; i = next.fetch_add(1)) {
            lqm_hf_problem& p = problems[order[i]];
            p.converged = 0;
            p.status = lqm_hartreefock_scf(wc, p.N, d, p.nRh, p.Rh, p.Th, p.nRv, p.Rv, p.Tv, p.nRr, p.Rr, p.rho,
                                           p.Nk, p.ks, p.kweights, p.filling, T, beta, tol, max_iter, flags,
                                           p.result, &p.converged, nullptr, nullptr, nullptr, nullptr);
            if (p.status && errst[w] == LQM_OK) {
                errst[w] = p.status;
                errs[w] = g_err;
            }
            if (ctx->progress) {
                std::lock_guard<std::mutex> lk(prog_mu);
                ctx->progress(order[i], p.status ? -1 : (int)p.result[7], p.converged, ctx->progress_user);
            }
        }
        lqm_context_destroy(wc);
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < W; ++w) pool.emplace_back(worker, w);
    worker(0);
    for (auto& t : pool) t.join();
    for (int w = 0; w < W; ++w)
        if (errst[w] != LQM_OK) return fail((lqm_status)errst[w], "batch: " + errs[w]);
    return LQM_OK;
}

}  // extern "C"
