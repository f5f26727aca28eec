// Internal declarations shared by the HIP translation units of libexafac_hip.so.
// Nothing here crosses the C-ABI (include/exafac_hip.h).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>
#include <stdexcept>

namespace exafac {

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define EX_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            throw ::exafac::Error(2, std::string(#call) + ": " + hipGetErrorString(e_) \
                                         + " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define EX_CHECK(cond, code, msg)                         \
    do {                                                  \
        if (!(cond)) throw ::exafac::Error(code, (msg));  \
    } while (0)

constexpr int kMaxRank = 64;
constexpr int kMaxModes = 8;

// ---------------------------------------------------------------- device memory
// Owning device buffer (bytes), grown on demand; never shrinks.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; bytes = 0;
    }
    void ensure(size_t n) {
        if (n <= bytes) return;
        // queued kernels may still read the old allocation: drain the device
        // before freeing it (buffers only grow, so this is rare)
        if (p) (void)hipDeviceSynchronize();
        release();
        size_t want = n < 256 ? 256 : n;
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error(5, "hipMalloc(" + std::to_string(want) + " bytes) failed: " + hipGetErrorString(e));
        }
        bytes = want;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Packed upper triangle of an R x R symmetric matrix, row-major over the
// upper triangle: (a, b), a <= b, at a*R - a*(a-1)/2 + (b - a).
__host__ __device__ inline int tri_size(int R) { return R * (R + 1) / 2; }
__host__ __device__ inline int tri_index(int R, int a, int b) { return a * R - (a * (a - 1)) / 2 + (b - a); }
// Node stride in elements: packed size rounded up to 16 elements (128 B in FP64).
__host__ __device__ inline int tri_stride(int R) { return (tri_size(R) + 15) & ~15; }
// FP32 node stride: packed size rounded up to 32 elements (128 B).
__host__ __device__ inline int tri_stride32(int R) { return (tri_size(R) + 31) & ~31; }
// Zeroed bytes after each STS-CP row-tree and eigen-tree node array: the draw
// reads all NV W-wide vectors of a node's G lanes (up to 17 KB) without a
// bound check, so the reads past the last node land here.
constexpr size_t kNodeSlackBytes = 32768;
// Packed size rounded up to a multiple of 64 (a weight table read G lanes at a
// time, G <= 64, with no bound check).
__host__ __device__ inline int tri_pad64(int R) { return (tri_size(R) + 63) & ~63; }

// Heap-ordered segment tree over ceil(n/F) leaves (PartitionTree geometry,
// partition_tree.hpp:82-115).
struct TreeGeom {
    int64_t n = 0, F = 0;
    uint32_t leaf_count = 0, node_count = 0, lfill_level = 0, lfill_count = 0;
    uint32_t nodes_upto_lfill = 0, nodes_before_lfill = 0, complete_level_offset = 0;
    void init(int64_t n_, int64_t F_) {
        n = n_; F = F_;
        leaf_count = (uint32_t)((n + F - 1) / F);
        node_count = 2 * leaf_count - 1;
        lfill_level = 0; lfill_count = 1;
        while (lfill_count * 2 <= leaf_count) { lfill_level++; lfill_count *= 2; }
        nodes_upto_lfill = lfill_count * 2 - 1;
        nodes_before_lfill = lfill_count - 1;
        complete_level_offset = nodes_before_lfill - (node_count - nodes_upto_lfill) / 2;
    }
    __host__ __device__ int64_t leaf_idx(int64_t c) const {
        return c >= (int64_t)nodes_upto_lfill ? c - nodes_upto_lfill : c - complete_level_offset;
    }
    __host__ __device__ bool is_leaf(int64_t c) const { return 2 * c + 1 >= (int64_t)node_count; }
};

// Bit layout of the fiber key: coordinates of every mode except `skip`, most
// significant first in ascending mode order.
struct KeyLayout {
    int N = 0;
    int shift[kMaxModes];
    int bits = 0;
};
KeyLayout make_key_layout(int N, int skip, const int64_t* dims);

// Per-mode lookup for the sampled MTTKRP.  A fiber is the set of nonzeros
// sharing every coordinate but the target mode's; fibers are numbered in
// ascending key order.  The nonzeros are kept in target-row order (ties in
// fiber order) with their fiber numbers, so one pass over them, keeping those
// whose fiber is sampled, produces each output row in a fixed order.
struct ModeLookup {
    int mode = -1;
    KeyLayout kl;
    int64_t nfib = 0;
    int64_t nnz = 0;
    DevBuf fkey;    // u64[nfib] distinct fiber keys, ascending
    DevBuf rrow;    // u32[nnz] target-mode row, ascending
    DevBuf rfib;    // u32[nnz] fiber number
    DevBuf rval;    // f64[nnz]
    DevBuf rval32;  // f32[nnz], built on first use by the mixed-precision path
    DevBuf cmap;    // u64[nfib]: 2^32 | design row of the fibers sampled by the running call, else 0
    DevBuf filt;    // u32[2^14]: bit (f mod 2^19) set for this step's sampled fibers f (zero between steps)
    // hit-list path
    DevBuf fptr;    // u32[nfib + 1]: first key-order position of each fiber
    DevBuf kpos;    // u32[nnz]: row-order position of each key-ordered nonzero
    DevBuf cpos;    // u32[nnz]: design row of the hit at a row-order position (valid where its bit is set)
    DevBuf marks;   // u8, whole 4096-position chunks: 1 at this step's hit positions (zero between steps)
    DevBuf cpart;   // u64 per-workgroup counter partials of the passes
    // bucketed path (built on first use): key-order target row and value of each nonzero
    DevBuf krow;    // u32[nnz]
    DevBuf kval;    // f64[nnz]
    DevBuf kval32;  // f32[nnz]
    int64_t maxfib = 0;    // nonzeros in the longest fiber
    int mttkrp_path = 0;   // 0: chosen per step from the sizes, 1: full row pass, 2: hit bitmap, 3: bucketed hits
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipDeviceProp_t prop;
    DevBuf scratch[8];
};

// ---------------------------------------------------------------- kernels (launchers)
// All launchers enqueue on `s` and never synchronise.
namespace k {
// dense
void gram(hipStream_t s, const double* A, int64_t m, int R, double* G, DevBuf& part,
          const uint64_t* m_dev = nullptr);
// G = A^T A with the block partials in `part` (part_bytes; no allocation)
void gram_into(hipStream_t s, const double* A, int64_t m, int R, double* G, double* part, size_t part_bytes);
// G = A^T A for FP32 A with FP64 products and sums
void gram_f32in(hipStream_t s, const float* A, int64_t m, int R, double* G, DevBuf& part,
                const uint64_t* m_dev = nullptr);
void eigh(hipStream_t s, const double* A, int R, double* lam, double* V, int ncopies = 1);
// only_if: when given, the launch does nothing unless *only_if != 0.
// relative: eigenvalues <= tol * (largest diagonal entry of M) are dropped
// instead of eigenvalues <= tol.
void pinv(hipStream_t s, const double* M, int R, double tol, double* out, int ncopies = 1,
          const int* only_if = nullptr, bool relative = false);
void dagat(hipStream_t s, const double* A, int64_t m, int R, const double* G, double* res);
void dsymm(hipStream_t s, const double* G, const double* Y, int64_t m, int R, double* out);
void col_sumsq(hipStream_t s, const double* U, int64_t m, int R, double* sigma, DevBuf& part);
void scale_cols(hipStream_t s, double* U, int64_t m, int R, const double* sigma);
void cos_init(hipStream_t s, double* U, int64_t m, int R);
void chain_had_norm(hipStream_t s, const double* grams, int N, int R, const double* sigma, double* out);
// out = elementwise product of grams[k] (k != exclude), k ascending
void chain_had(hipStream_t s, const double* grams, int N, int exclude, int R, double* out);
// preprocessing / norms
void preprocess(hipStream_t s, double* vals, int64_t nnz, int kind);
void sumsq(hipStream_t s, const double* v, int64_t n, double* out, DevBuf& part);
void plain_sum(hipStream_t s, const double* v, int64_t n, double* out, DevBuf& part);
// ARLS
void arls_cdf(hipStream_t s, const double* scores, int64_t I, double* cdf, double* sum, DevBuf& part);
void arls_draw(hipStream_t s, int N, int j, int64_t J, const double* const* cdfs, const double* const* scores,
               const double* sums, const int64_t* dims, uint64_t seed, uint64_t sweep,
               const double* draws, uint32_t* samples, double* weights, const uint64_t* sweep_dev = nullptr);
// STS
void tree_build(hipStream_t s, const double* U, const TreeGeom& g, int R, double* nodes);
// internal nodes of a tree whose leaves are built (node = left + right)
template <typename T>
void level_sums(hipStream_t s, const TreeGeom& g, int S, T* nodes);
// the root node (packed upper, the Gram of the whole factor) as a full R x R FP64 matrix
template <typename T>
void unpack_root(hipStream_t s, const T* nodes, int R, double* G);
void sts_compute_m(hipStream_t s, int N, int j, int R, const double* grams, double* M, double* lam);
void eig_tree_build(hipStream_t s, int N, int j, int R, const double* M, const double* grams, double* eig_nodes);
void eig_tree_build32(hipStream_t s, int N, int j, int R, const double* M, const double* grams, float* eig_nodes);
struct StsModeArgs {
    const void* rows;      // factor U_k (I_k x R), double or float as the nodes
    const void* nodes;     // packed row-tree nodes (stride tri_stride / tri_stride32)
    const void* eig;       // packed eigen-tree nodes (2R-1), same precision and stride as the row tree
    const double* w;       // R x R scaled eigenvectors, rows = components
    TreeGeom g;
    TreeGeom ge;
};
// node_bytes = 8: FP64 rows and row-tree nodes; 4: FP32 (the mixed path).
// first_cdf (FP32 only): the prefix sums of sts_first_cdf for the first
// non-target mode, whose draws then search them instead of descending the
// trees (nullptr: every mode descends).
void sts_draw(hipStream_t s, int N, int j, int R, int64_t J, const StsModeArgs* modes,
              const double* pinv_packed, uint64_t seed, uint64_t sweep,
              const double* draws_row, const double* draws_eig, uint32_t* samples, double* weights,
              int node_bytes = 8, const uint64_t* sweep_dev = nullptr,
              unsigned long long* dropped = nullptr, const double* first_cdf = nullptr);
// The first non-target mode's draw distribution in closed form: with the
// running product h = 1 the row query is an eigen row w_v, so the leaf
// masses m_v(l) = w_v^T G_l w_v (FP32 row-tree leaf Grams) give, per
// component v, prefix sums over the leaves in two levels (FP64): within
// each tile of 32 leaves, cdf[v L + l], and over the NT = ceil(L / 32) tile
// totals, cdf[R L + v NT + q]; the last is v's total.  R <= 64; cdf holds
// R (L + NT) doubles.
void sts_first_cdf(hipStream_t s, const StsModeArgs& m, int R, double* cdf);
// packed[e] = (1 on the diagonal, 2 off it) x M[a][b] for the packed upper
// element e = (a, b), zeros up to tri_pad64(R) entries (the draw's weight table).
void pack_upper(hipStream_t s, const double* M, int R, double* packed);
// sparse
void keys_from_coords(hipStream_t s, const uint32_t* coords, int64_t n, const KeyLayout& kl, uint64_t* keys,
                      uint32_t* iota);
// Builds L (fkey, rrow, rfib, rval, bitmap, cmap) from the COO tensor.
void build_lookup(hipStream_t s, const uint32_t* coords, const double* vals, int64_t nnz, int N, int mode,
                  const int64_t* dims, ModeLookup& L, DevBuf* tmp);
// Returns J_u, or -1 when Ju_dev is given: the count then stays on the device
// (Ju_dev) with no host synchronisation, and is added to *Ju_total if given.
// Kernels taking (bound, *_dev) row counts use min(bound, *dev).
int64_t dedup(hipStream_t s, const uint32_t* samples, const double* weights, int64_t J, int N, const KeyLayout& kl,
              uint32_t* sd, double* wd, uint64_t* kd, DevBuf* tmp, uint64_t* Ju_dev = nullptr,
              uint64_t* Ju_total = nullptr);
// The same unique tuples, keys and weight x multiplicity by hashing (keys of
// at most 63 bits), in the order of first occurrence among the samples.
// table: the hash table (kept zeroed between calls); work: per-sample slots.
int64_t dedup_hash(hipStream_t s, const uint32_t* samples, const double* weights, int64_t J, int N,
                   const KeyLayout& kl, uint32_t* sd, double* wd, uint64_t* kd, DevBuf& table, DevBuf& work,
                   uint64_t* Ju_dev = nullptr, uint64_t* Ju_total = nullptr);
void design(hipStream_t s, const uint32_t* sd, const double* wd, int64_t Ju, int N, int j, int R,
            const double* const* U, double* d_gram, double* d_mttkrp, const uint64_t* Ju_dev = nullptr);
// Y (I_j x R, zero-filled here unless zero_y is false: then the caller has
// cleared it) = sampled MTTKRP of the Ju deduplicated samples with fiber keys
// kd and design rows `design`.  V = double (FP64 path) or float (mixed path):
// values (L.rval / L.rval32), design rows and Y.  counts (optional, device
// u64[2]): the hits and the touched rows are added.
template <typename V>
void sampled_mttkrp(hipStream_t s, ModeLookup& L, const V* vals, const uint64_t* kd, int64_t Ju, const V* design,
                    int R, V* Y, int64_t I_j, DevBuf* tmp, uint64_t* counts = nullptr,
                    hipEvent_t* scan_events = nullptr, const uint64_t* Ju_dev = nullptr, bool zero_y = true);
// p[0, bytes) = 0 by a fill kernel (bytes a multiple of 4; no memset node
// under graph capture).
void zero_bytes(hipStream_t s, void* p, size_t bytes);
// Y (I_j x R, zero-filled here) = exact MTTKRP: every nonzero x at (i_0..i_N)
// adds x * prod_{k != mode} U_k[i_k] to row i_mode.  U: the N factors in V.
template <typename V>
void exact_mttkrp(hipStream_t s, ModeLookup& L, const V* vals, const V* const* U, const int64_t* dims, int R,
                  V* Y, int64_t I_j, DevBuf* tmp, uint64_t* counts = nullptr);
// Adjoint of the sampled MTTKRP: dD = per-sample fiber sums of x * G[row]
// (Ju x R, written), then grads[k] (k != j, non-null) += dD * dDesign/dU_k.
void smttkrp_adjoint(hipStream_t s, const ModeLookup& L, const uint64_t* kd, const uint32_t* sd, const double* wd,
                     int64_t Ju, int N, int j, int R, const double* const* U, const double* G,
                     double* const* grads, double* dD);
// out[0..n) = NaN unless *count == want: a device-side argument check whose
// failure the caller sees in the output, with nothing waited for.
void poison_unless(hipStream_t s, const uint64_t* count, int64_t want, double* out, int64_t n);
void fit_bmb(hipStream_t s, const uint32_t* coords, const double* vals, int64_t nnz, int N, const double* const* U,
             const double* sigma, int R, double* out, DevBuf& part);
// final stages shared with the mixed-precision kernels
void gram_final_launch(hipStream_t s, const double* part, int nb, int R, double* G);
void col_sumsq_final_launch(hipStream_t s, const double* part, int nb, int R, double* sigma);
// mixed precision (FP32 data, BF16 MFMA Grams, FP64 R x R solves)
void gram_bf16(hipStream_t s, const float* A, int64_t m, int R, double* G, DevBuf& part,
               const uint64_t* m_dev = nullptr);
void tree_build_f32(hipStream_t s, const float* U, const TreeGeom& g, int R, float* nodes);
// Gauss-Jordan inverse of an SPD matrix; pivots <= tol_rel * max diagonal are
// dropped (their row and column zeroed) and counted into *dropped if given.
void spd_inverse(hipStream_t s, const double* M, int R, double tol_rel, double* out, int* dropped = nullptr);
// The pseudo-inverse: spd_inverse, replaced when a pivot was dropped (M
// numerically singular) by the Moore-Penrose pseudo-inverse -- for R <= 32 a
// pivoted Cholesky in the same workgroup (same relative tolerance), above 32
// the Jacobi pinv as a second launch (tol_rel as its absolute tolerance).
// flag: one device int, the dropped-pivot count (read by that second launch).
void spd_pinv(hipStream_t s, const double* M, int R, double tol_rel, double* out, int* flag);
void sts_compute_m_mixed_launch(hipStream_t s, int N, int j, int R, const double* grams, double* M);
void to_f32(hipStream_t s, const double* a, int64_t n, float* b);
void to_f64(hipStream_t s, const float* a, int64_t n, double* b);
void design32(hipStream_t s, const uint32_t* sd, const double* wd, int64_t Ju, int N, int j, int R,
              const float* const* U, float* dg, float* dm, const uint64_t* Ju_dev = nullptr);
void dsymm32(hipStream_t s, const float* G, const float* Y, int64_t m, int R, float* out);
}  // namespace k
// FROSTT .tns ingest (host): line count and mode count, then the raw
// coordinates (nnz x N, as written), values and per-mode minimum / maximum.
void tns_scan(const char* path, int64_t& nnz, int& N);
void tns_read(const char* path, int64_t nnz, int N, uint64_t* coords, double* vals, uint64_t* mins, uint64_t* maxs);
namespace k {
void renormalize32(hipStream_t s, float* U, int64_t m, int R, double* sigma, DevBuf& part);
// sigma_c = ||(Y Gi)[:, c]|| from GY = Y^T Y; Ms = Gi diag(1 / sigma) (FP32);
// gu (optional, R x R): U^T U for U = Y Ms, as Ms^T GY Ms in FP64
void sigma_from_gram(hipStream_t s, const double* GY, const double* Gi, int R, double* sigma, float* Ms,
                     double* gu = nullptr);
void dagat32(hipStream_t s, const float* A, int64_t m, int R, const double* G, double* res);
// sort / scan primitives (rocPRIM)
size_t sort_pairs_u64_u32(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* kin, uint64_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int bits);
size_t sort_pairs_u32_u64(hipStream_t s, void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint64_t* vin, uint64_t* vout, int64_t n, int bits);
size_t sort_pairs_u32_u32(hipStream_t s, void* temp, size_t temp_bytes, const uint32_t* kin, uint32_t* kout,
                          const uint32_t* vin, uint32_t* vout, int64_t n, int bits);
size_t exclusive_scan_u32(hipStream_t s, void* temp, size_t temp_bytes, const uint32_t* in, uint32_t* out, int64_t n);
size_t exclusive_scan_u64(hipStream_t s, void* temp, size_t temp_bytes, const uint64_t* in, uint64_t* out, int64_t n);
}  // namespace k

// ---------------------------------------------------------------- launch variants
// Compiled alternatives an offline autotuner chooses between
// (rdist_tensor_amd/autotune.py measures them and writes tuning.txt beside
// the library, which the first context loads).  Value 0 is the built-in
// default; every variant computes the same quantities.
enum TuneKnob {
    kTuneDrawF64R16,   // STS-CP draw, FP64 nodes, R <= 16: (lanes per sample, vectors per lane, threads)
    kTuneDrawF64R32,   // FP64, 16 < R <= 32
    kTuneDrawF32R16,   // FP32 nodes, R <= 16
    kTuneDrawF32R32,   // FP32, 16 < R <= 32
    kTuneDrawF32R64,   // FP32, 45 < R <= 64
    kTuneScanF64,      // sampled-MTTKRP row pass, FP64: lanes per hit x 1 or x 2
    kTuneScanF32,      // FP32
    kTuneFirstF32,     // FP32 STS-CP draw, first non-target mode: prefix sums (0) or tree descents (1)
    kTuneKnobs
};
struct TuneInfo {
    const char* name;
    int variants;
};
const TuneInfo& tune_info(int knob);
int tune_value(int knob);
void tune_select(int knob, int value);

// ---------------------------------------------------------------- batched sweeps
// Device descriptors of the batched CP-ARLS-LEV sweep (one workgroup runs
// whole tensors).  V = float (mixed precision) or double (FP64) for factor
// rows, nonzero values, design rows and MTTKRP rows.
constexpr int kBatchMaxRank = 32;
constexpr int kBatchMaxJ = 8192;

struct BatchMode {
    void* U;                  // V[I x R]: the factor
    int64_t I;
    // lookup with this mode as the target (ModeLookup)
    const uint64_t* fkey;     // nfib distinct fiber keys, ascending
    int64_t nfib;
    const uint32_t* fptr;     // nfib + 1 key-order offsets
    const uint32_t* kpos;     // row-order position of each key-ordered nonzero
    const uint32_t* rrow;     // row of each row-order position
    const void* rval;         // V value of each row-order position
    uint32_t* rowptr;         // I + 1: first row-order position of each row
    int shift[kMaxModes];     // fiber key layout (-1 for this mode)
    // leverage state over the rows where the factor is nonzero
    uint32_t* list;           // rows, ascending (capacity I)
    double* csc;              // their leverage scores
    double* ccdf;             // CDF over them (last entry 1)
};

struct BatchTensor {
    int N, R;
    int64_t nnz;
    uint64_t seed;
    BatchMode m[kMaxModes];
    double* sigma;            // R: column norms of the last updated factor
    uint32_t* cnt;            // [kMaxModes] listed rows per mode
    double* ssum;             // [kMaxModes] score sums
    unsigned long long* stats;  // running totals: [0] nonzeros iterated, [1] unique samples, [2] rows updated
    // scratch
    uint32_t* samples;        // J x N
    double* weights;          // J
    uint32_t* sd;             // J x N unique tuples, first-occurrence order
    uint64_t* kd;             // J their fiber keys
    double* wd;               // J weight x multiplicity
    void* dg;                 // V[J x R] design rows (Gram)
    void* dm;                 // V[J x R] design rows (MTTKRP)
    void* Yc;                 // V[Imax x R] MTTKRP rows of the listed rows
    uint32_t* hu;             // nnz: design row of a hit position
    uint32_t* hstamp;         // nnz: step stamp of a hit position
    uint32_t* rstamp;         // Imax: step stamp of a hit row
};

// Read by every workgroup at launch start; the last workgroup to finish
// advances sweep and epoch and resets the queue.
struct BatchCtl {
    unsigned long long sweep;  // keys the draws of the launch
    unsigned int epoch;        // launches since creation (step stamps)
    unsigned int next;         // tensor queue
    unsigned int done;         // finished workgroups
    unsigned int pad;
};

namespace k {
// list = iota and rowptr for every mode of every tensor
void batch_init(hipStream_t s, const BatchTensor* ts, int n);
// one sweep of every tensor (grid workgroups); vbytes 4 (mixed) or 8 (FP64)
void batch_sweep(hipStream_t s, const BatchTensor* ts, int n, int64_t J, BatchCtl* ctl, int grid, int vbytes);
}  // namespace k

inline unsigned grid_for(int64_t work, int per_block, int64_t cap = 1 << 20) {
    int64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

}  // namespace exafac
