// liblensim: HIP kernels for gfx950 and the C ABI declared in include/lensim.h.
//
// Pipeline of one batched image / likelihood call (all on one stream):
//   prep_psf_kernel  kwargs slots -> per-(image, profile) records (+ EPL series
//                    coefficients), float64 and float32 copies; in the same
//                    launch the PSF preparation: the flipped kernel for the
//                    direct convolution, or the fp16 hi/lo B fragments of the
//                    matrix-core convolution
//   render_kernel /  one thread per pixel: s x s sub-rays generated from the
//   render_fast_...  grid description, ray shooting through every deflector,
//                    source light at the deflected position and lens light at
//                    the ray, mean over the sub-pixel block, x pixel_width^2 x
//                    flux_scaling (the fast kernel for compile-time profile
//                    lists of the common model family)
//   conv_kernel /    'same' convolution (zero or edge padding): LDS-tiled
//   conv_mfma_kernel direct stencil in the compute type, or fp16 hi/lo split
//                    products on the matrix cores (mixed form); optionally
//                    fused with the per-pixel chi^2 and a block reduction
//   logl_kernel      fixed-order sum of the per-tile chi^2 partials
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>

#include "grad.hpp"
#include "profiles.hpp"

using namespace lensim;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int hip_fail(hipError_t e, const char *where) {
  return fail(LENSIM_EDEVICE, "%s: %s", where, hipGetErrorString(e));
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// ---- bounds-checked build (-DLENSIM_BOUNDS_CHECK, liblensim_checked.so) ----
// Each entry point registers the buffers of its call (workspace, parameters,
// data, PSF, outputs) before its launches.  A checked access AT(p, i) tests
// that element i of p lies inside one of them; one that does not is diverted
// to a private sink and the first such access's source line is recorded, so
// the kernel touches no memory outside the call's buffers and still ends
// normally.  After the call the entry point waits for its stream and returns
// LENSIM_EDEVICE naming that line, and clears the ranges again: a call that
// registers none (graph capture, which cannot wait for its stream) runs
// unchecked.  In the normal build AT(p, i) is p[i].
constexpr int kMaxRanges = 24;
struct CheckRanges {
  int n;
  unsigned long long lo[kMaxRanges], hi[kMaxRanges];
};
#ifdef LENSIM_BOUNDS_CHECK
__constant__ CheckRanges g_ranges;
__device__ unsigned int g_violation;
__device__ __align__(16) unsigned char g_sink[64];

__device__ __forceinline__ bool in_ranges(const void *p, size_t n) {
  const unsigned long long a = reinterpret_cast<unsigned long long>(p);
  if (g_ranges.n == 0) return true;  // no ranges registered: unchecked
  for (int r = 0; r < g_ranges.n; ++r)
    if (a >= g_ranges.lo[r] && a + n <= g_ranges.hi[r]) return true;
  return false;
}

__device__ __noinline__ void report_violation(int line) { atomicCAS(&g_violation, 0u, unsigned(line)); }

template <typename T>
__device__ __forceinline__ T &checked_at(T *p, size_t i, int line) {
  if (!in_ranges(p + i, sizeof(T))) {
    report_violation(line);
    return *reinterpret_cast<T *>(g_sink);
  }
  return p[i];
}
#define AT(p, i) checked_at((p), size_t(i), __LINE__)
// a base pointer and extent the kernel reads from with fixed offsets (records)
#define CHECK_SPAN(p, n) \
  do {                   \
    if (!in_ranges((p), (n))) report_violation(__LINE__); \
  } while (0)
#else
#define AT(p, i) (p)[i]
#define CHECK_SPAN(p, n) \
  do {                   \
  } while (0)
#endif

// while a call is captured into a graph (lensim_graph_capture_loglike) no
// timing events are recorded (they would become nodes of the graph) and the
// checked build registers no ranges (it would have to wait for the stream)
thread_local bool g_capturing = false;

// The buffers of one call, for the checked build (no-ops otherwise).
class CallRanges {
 public:
  void add(const void *p, size_t bytes) {
    if (!p || !bytes) return;
    if (r_.n >= kMaxRanges) {
      overflow_ = true;
      return;
    }
    r_.lo[r_.n] = reinterpret_cast<unsigned long long>(p);
    r_.hi[r_.n] = r_.lo[r_.n] + bytes;
    ++r_.n;
  }
#ifdef LENSIM_BOUNDS_CHECK
  int begin(hipStream_t st) {
    if (g_capturing) return LENSIM_OK;
    if (overflow_) return fail(LENSIM_EINVAL, "bounds check: more than %d buffers in one call", kMaxRanges);
    const unsigned int zero = 0;
    hipError_t e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_ranges), &r_, sizeof(r_), 0, hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
      e = hipMemcpyToSymbolAsync(HIP_SYMBOL(g_violation), &zero, sizeof(zero), 0, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);  // the host copies above read r_ and zero
    return e == hipSuccess ? LENSIM_OK : hip_fail(e, "bounds check set-up");
  }
  int end(hipStream_t st, int rc) {
    if (g_capturing) return rc;
    unsigned int line = 0;
    const int none = 0;
    hipError_t e = hipStreamSynchronize(st);
    if (e == hipSuccess) e = hipMemcpyFromSymbol(&line, HIP_SYMBOL(g_violation), sizeof(line));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(g_ranges), &none, sizeof(none));
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "bounds check");
    if (line)
      return fail(LENSIM_EDEVICE, "bounds check: an access at lensim.hip:%u fell outside the call's buffers", line);
    return LENSIM_OK;
  }
#else
  int begin(hipStream_t) { return LENSIM_OK; }
  int end(hipStream_t, int rc) { return rc; }
#endif

 private:
  CheckRanges r_{};
  bool overflow_ = false;
};

constexpr int kWin = 8;  // register window (outputs per thread) of the convolution
constexpr int kGradPix = 4;  // pixels per thread of the gradient kernel
constexpr size_t kMaxLds = 160 * 1024;  // LDS per workgroup on gfx950

// ---- optional kernel timing (lensim_timing_*) ----
struct TimedCall {
  hipEvent_t ev[3];  // before render, after render / before conv, after conv
};
bool g_timing = false;
TimedCall *g_calls = nullptr;
int g_ncalls = 0, g_capcalls = 0;

hipEvent_t timing_event(int slot) {
  if (!g_timing) return nullptr;
  if (slot == 0) {
    if (g_ncalls == g_capcalls) {
      const int cap = g_capcalls ? 2 * g_capcalls : 256;
      TimedCall *n = static_cast<TimedCall *>(realloc(g_calls, sizeof(TimedCall) * cap));
      if (!n) return nullptr;
      for (int i = g_capcalls; i < cap; ++i)
        for (int k = 0; k < 3; ++k) n[i].ev[k] = nullptr;
      g_calls = n;
      g_capcalls = cap;
    }
    ++g_ncalls;
  }
  TimedCall &c = g_calls[g_ncalls - 1];
  if (!c.ev[slot] && hipEventCreate(&c.ev[slot]) != hipSuccess) c.ev[slot] = nullptr;
  return c.ev[slot];
}


void timing_mark(int slot, hipStream_t st) {
  if (!g_timing || g_capturing) return;
  hipEvent_t e = timing_event(slot);
  if (e) (void)hipEventRecord(e, st);
}

// Record index space: [0, n_lens) inline deflectors, [n_lens, n_defl) bulk
// deflectors, [n_defl, n_defl + n_source) source light, then lens light.
// types[] holds the inline lens, source and lens-light types (no bulk).
struct ModelArgs {
  int32_t n_lens, n_bulk, n_defl, n_source, n_light, n_prof, n_epl, has_halo, bulk_type;
  int32_t has_nie_nfw;  // an SIE / NIE or NFW deflector (inline or bulk), or halos on several planes: the float
                        // form deflects SIE / NIE and the NFW family in double
  SubSteps sub;         // sub-ray steps of the call's grid, for the records (with_grid); zero without a grid
  const int32_t *bulk_types;
  int32_t types[LENSIM_MAX_PROFILES];
  int32_t epl_index[LENSIM_MAX_PROFILES];
  double alpha_scaling, flux_scaling;
  // multi-plane ray tracing (mp_n == 0: single plane)
  int32_t mp_n;
  const int32_t *mp_order;
  const double *mp_tij, *mp_tz, *mp_r2p;
  int64_t mp_stride;
};

__host__ __device__ inline int type_of(const ModelArgs &m, int k) {
  if (k < m.n_lens) return m.types[k];
  if (k < m.n_defl) return m.bulk_type ? m.bulk_type : m.bulk_types[k - m.n_lens];
  return m.types[k - m.n_bulk];
}

struct Layout {
  size_t recd, recf, coefd, coeff, dcoefd, dcoeff, psf, lowres, partial, blkmax, bfrag;
  size_t gmodel, glow, gpart;  // gradient: dlogL/dmodel, dlogL/dlowres, per-block partial gradients
  size_t gbeta;                // gradient: dS/dbeta of the sources at every ray (0: none)
  size_t gpart_mp;             // gradient, several planes: per-block partials of the deflectors (0: none)
  size_t gpad, gext;           // edge-padded PSF adjoint: padded dlogL/dmodel and its correlation
  size_t gchr, glu;            // supersampled convolution adjoint buffers
  size_t hr, chr, lowu, lowc, psf_low;  // supersampled convolution: ray fluxes, their convolution,
                                        // pooled unconvolved image and its low-resolution convolution
  size_t total;
  int n_tiles;
};

// Convolution tile shape (outputs per block): TX columns, TY rows; each
// thread owns kWin consecutive rows of one column, so a block has
// TX * TY / kWin threads and lanes of a wave read consecutive LDS words.
template <typename T>
struct ConvTile {
  static constexpr int TY = 32, TX = 32;  // 128 threads
};

// kernel height padded to a multiple of the register window
int kyp_of(int ky) { return (ky + kWin - 1) / kWin * kWin; }

// ---- MFMA convolution geometry (mixed form) ----
// Each kernel row a contributes D(16x16) += A(16x32k) B(32k x16) per output
// tile, with A = 16 image rows of the padded tile and B[k][n] = Kf[a][k - n]
// (the band of kernel row a).  nks K-steps of 32 cover 16 + kx - 1 columns.
constexpr int kMfWaves = 4;  // waves per block, each 16 output rows
constexpr int kMfTY = 16 * kMfWaves;
// 16-column tiles per wave: TL in [2, 4], chosen per call (mfma_plan)
int mfma_nks(int kx) { return (16 + kx - 1 + 31) / 32; }
constexpr int kMfMaxNks = 4;  // kernels up to 113 columns
// Kernel rows per unrolled loop body (and per float -> double fold): the B
// fragments of that many rows (about 4 K-steps) are in flight at once.
constexpr int mfma_unroll(int nks) { return (4 + nks - 1) / nks; }
// B-fragment steps: ky padded to a multiple of the unroll with zero rows.
int mfma_steps(int kx, int ky) {
  const int nks = mfma_nks(kx), u = mfma_unroll(nks);
  return (ky + u - 1) / u * u * nks;
}
int mfma_tw(int kx, int tl) { return 16 * (tl - 1) + 32 * mfma_nks(kx); }
// LDS row stride in halves.  The A reads (ds_read_b128, lane l at row l % 16,
// 16-byte column l / 16) are free of bank conflicts when the stride in dwords
// is 8 (mod 16), as measured on gfx950 (strides of 4 or 12 (mod 16) dwords
// conflict, 0 (mod 64) serialises).  tw = 16 (tl - 1) + 32 nks halves: for
// even tl it is 16 (mod 32) and unpadded; for tl = 3 it is padded by 16.
int mfma_ls(int kx, int tl) {
  const int tw = mfma_tw(kx, tl);
  return tw % 32 == 16 ? tw : tw + 16;
}
size_t mfma_smem(int kx, int ky, int tl) {
  return size_t(kMfTY + ky - 1) * mfma_ls(kx, tl) * 2 * sizeof(_Float16);
}

// ---- float64 matrix-core convolution (FP64 form) ----
// v_mfma_f64_16x16x4_f64 (lane l holds A[l % 16][l / 16] and B[l / 16][l % 16]
// of a K-step): per kernel row a, D(16 x 16) += A(16 x 4 nk4) B(4 nk4 x 16)
// with A = 16 image rows of the padded tile and B[k][n] = Kf[a][k - n]; nk4
// K-steps of 4 cover 16 + kx - 1 columns.  The LDS row stride is 2 (mod 16)
// doubles: the 32 lanes of each half of a ds_read_b64 (16 rows x 2 columns)
// then fall on distinct banks.
int mf64_nk4(int kx) { return (16 + kx - 1 + 3) / 4; }
int mf64_tw(int kx, int tl) { return 16 * (tl - 1) + 4 * mf64_nk4(kx); }
int mf64_ls(int kx, int tl) {
  const int tw = mf64_tw(kx, tl);
  return tw + ((2 - tw % 16) + 16) % 16;
}
size_t mf64_smem(int kx, int ky, int tl) { return size_t(kMfTY + ky - 1) * mf64_ls(kx, tl) * sizeof(double); }
// K-steps (kernel rows x nk4) padded with zero fragments to a multiple of the
// kernel's prefetch depth, so its step loop has no partial group
constexpr int kMf64B = 8;
int mf64_steps(int kx, int ky) { return (ky * mf64_nk4(kx) + kMf64B - 1) / kMf64B * kMf64B; }

// Blocks of 64 rows x 16 tl columns.  All of a call's blocks are resident at
// once on this workload's sizes, so the time is that of the busiest CU:
// ceil(blocks / CUs) blocks of tl wave-tiles.  tl is the one in [2, 4] with
// the least such load (ties to the wider tile, which re-reads fewer A
// fragments): for 180^2 x 64 images, 768 blocks of 48 columns (3 per CU)
// against 576 of 64 (2.25 per CU, so 3 on some).
struct MfmaPlan {
  int tl, n_tiles_x, n_tiles;
};
int cu_count() {
  static int n = 0;
  if (n == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      n = v;
    else
      n = 256;
  }
  return n;
}
MfmaPlan mfma_plan(int nx, int ny, int n_img, int kx, int ky, bool f64 = false) {
  MfmaPlan best{0, 0, 0};
  long best_load = 0;
  const int cus = cu_count();
  for (int tl = 4; tl >= 2; --tl) {
    if ((f64 ? mf64_smem(kx, ky, tl) : mfma_smem(kx, ky, tl)) > kMaxLds) continue;
    const int ntx = (nx + 16 * tl - 1) / (16 * tl), nt = ntx * ((ny + kMfTY - 1) / kMfTY);
    const long blocks = long(nt) * n_img;
    const long load = (blocks + cus - 1) / cus * tl;
    if (best.tl == 0 || load < best_load) {
      best = MfmaPlan{tl, ntx, nt};
      best_load = load;
    }
  }
  return best;
}

// 16 x 16 output tiles of an image (the matrix-core convolution's chi^2 partials)
int n16_tiles(int nx, int ny) { return ((nx + 15) / 16) * ((ny + 15) / 16); }

template <typename T>
int n_tiles_of(int nx, int ny) {
  const int tx = (nx + ConvTile<T>::TX - 1) / ConvTile<T>::TX;
  const int ty = (ny + ConvTile<T>::TY - 1) / ConvTile<T>::TY;
  return tx * ty;
}

// Supersampled convolution plan (lensim_psf.supersampled on a grid with s > 1).
struct SSPlan {
  int on = 0;        // convolve on the supersampled grid
  int s = 1;
  int kxl = 0, kyl = 0;  // low-resolution split kernel (0: none)
};

Layout layout(const ModelArgs &m, int nx, int ny, int kx, int ky, int n_img, int precision,
              const SSPlan &ss = SSPlan()) {
  Layout L{};
  const size_t tsz = precision == LENSIM_MIXED ? 4 : 8;
  size_t off = 0;
  const size_t nrec = size_t(n_img) * m.n_prof * kRec;
  L.recd = off;
  off = align256(off + nrec * 8);
  L.recf = off;
  off = align256(off + nrec * 4);
  const size_t ncoef = size_t(n_img) * m.n_epl * LENSIM_EPL_NMAX;
  L.coefd = off;
  off = align256(off + ncoef * 8);
  L.coeff = off;
  off = align256(off + ncoef * 4);
  L.dcoefd = off;
  off = align256(off + ncoef * 8);
  L.dcoeff = off;
  off = align256(off + ncoef * 4);
  L.psf = off;
  off = align256(off + size_t(kx) * kyp_of(ky) * 8);
  L.lowres = off;
  off = align256(off + size_t(n_img) * nx * ny * tsz);
  L.n_tiles = precision == LENSIM_MIXED ? n_tiles_of<float>(nx, ny) : n_tiles_of<double>(nx, ny);
  const size_t npix = size_t(nx) * ny, npix_r = ss.on ? npix * ss.s * ss.s : npix;  // rendered pixels
  L.partial = off;  // per-tile (or per-block, or per 16 x 16 tile) chi^2 partials
  {
    size_t np_ = L.n_tiles;
    np_ = np_ > (npix + 255) / 256 ? np_ : (npix + 255) / 256;
    np_ = np_ > size_t(n16_tiles(nx, ny)) ? np_ : size_t(n16_tiles(nx, ny));
    off = align256(off + size_t(n_img) * np_ * 8);
  }
  L.blkmax = off;  // per render block max |value|, [n_img][ceil(rendered pixels / 256)]
  off = align256(off + size_t(n_img) * ((npix_r + 255) / 256) * 4);
  L.hr = L.chr = L.lowu = L.lowc = L.psf_low = off;
  if (ss.on) {
    L.hr = off;
    off = align256(off + size_t(n_img) * npix_r * tsz);
    L.chr = off;
    off = align256(off + size_t(n_img) * npix_r * tsz);
    if (ss.kxl > 0) {
      L.lowu = off;
      off = align256(off + size_t(n_img) * npix * tsz);
      L.lowc = off;
      off = align256(off + size_t(n_img) * npix * tsz);
      L.psf_low = off;
      off = align256(off + size_t(ss.kxl) * kyp_of(ss.kyl) * 8);
    }
  }
  L.bfrag = off;  // int exponent (16 B slot) then the MFMA B fragments (float64 fragments in the FP64 form)
  if (precision == LENSIM_MIXED) off = align256(off + 16 + size_t(mfma_steps(kx, ky)) * 2 * 64 * 16);
  else off = align256(off + size_t(mf64_steps(kx, ky)) * 64 * sizeof(double));
  L.total = off;
  return L;
}

// Layout of lensim_loglike_grad: the image layout plus the gradient buffers
// (on the ray grid when the convolution is supersampled).
// Blocks per image of the multi-plane gradient: one 256-ray sweep each where
// that keeps the whole launch near 2048 blocks, else several sweeps per block
// (so that the per-block partials of every deflector stay bounded).
int mp_grad_blocks(size_t rays, int n_img) {
  const long sweeps = long((rays + 255) / 256);
  const long want = (2048 + n_img - 1) / n_img;
  return int(sweeps < want ? sweeps : (want > 0 ? want : 1));
}

Layout grad_layout(const ModelArgs &m, int nx, int ny, int kx, int ky, int n_img, int precision, int edge,
                   const SSPlan &ss, int s) {
  Layout L = layout(m, nx, ny, kx, ky, n_img, precision, ss);
  const size_t tsz = precision == LENSIM_MIXED ? 4 : 8;
  const size_t npix = size_t(nx) * ny;
  const int cs = ss.on ? ss.s : 1;
  const size_t npix_r = npix * cs * cs;  // pixels of the grid the convolution runs on
  size_t off = L.total;
  L.gmodel = off;
  off = align256(off + size_t(n_img) * npix * tsz);
  L.glow = off;
  off = align256(off + size_t(n_img) * npix_r * tsz);
  L.gpart = off;
  off = align256(off + size_t(n_img) * m.n_prof * ((npix_r + 256 * kGradPix - 1) / (256 * kGradPix)) *
                           LENSIM_NPARAM * 8);
  L.gbeta = 0;
  const size_t rays = npix_r * (ss.on ? 1 : size_t(s) * s);
  if ((m.n_lens > 0 || m.mp_n > 0) && m.n_source > 0) {  // dS/dbeta or beta per ray (two values, compute type)
    L.gbeta = off;
    off = align256(off + size_t(n_img) * rays * 2 * tsz);
  }
  L.gpart_mp = 0;
  if (m.mp_n > 0 && m.n_defl > 0) {
    L.gpart_mp = off;
    off = align256(off + size_t(n_img) * m.n_defl * mp_grad_blocks(rays, n_img) * LENSIM_NPARAM * 8);
  }
  L.gpad = L.gext = L.gchr = L.glu = off;
  if (edge) {  // extended domain (ny cs + ky - 1) x (nx cs + kx - 1)
    const size_t next = size_t(nx * cs + kx - 1) * (ny * cs + ky - 1);
    L.gpad = off;
    off = align256(off + size_t(n_img) * next * tsz);
    L.gext = off;
    off = align256(off + size_t(n_img) * next * tsz);
  }
  if (ss.on) {
    L.gchr = off;  // dlogL / d(convolved rays)
    off = align256(off + size_t(n_img) * npix_r * tsz);
    L.glu = off;   // dlogL / d(pooled unconvolved rays), split kernel
    off = align256(off + size_t(n_img) * npix * tsz);
  }
  L.total = off;
  return L;
}

bool bulk_type_ok(int t) {
  return t == LENSIM_NULL || t == LENSIM_SIE || t == LENSIM_NIE || t == LENSIM_NFW || t == LENSIM_TNFW ||
         t == LENSIM_SHEAR || t == LENSIM_CONVERGENCE;
}

int model_args(const lensim_model *model, ModelArgs &m) {
  if (!model) return fail(LENSIM_EINVAL, "model is NULL");
  m = ModelArgs{};
  m.n_lens = model->n_lens;
  m.n_source = model->n_source;
  m.n_light = model->n_lens_light;
  m.n_bulk = model->n_bulk;
  if (m.n_lens < 0 || m.n_source < 0 || m.n_light < 0 || m.n_bulk < 0)
    return fail(LENSIM_EINVAL, "negative profile count");
  const int n_inline = m.n_lens + m.n_source + m.n_light;
  if (n_inline > LENSIM_MAX_PROFILES)
    return fail(LENSIM_EINVAL, "%d profiles exceed LENSIM_MAX_PROFILES=%d", n_inline, LENSIM_MAX_PROFILES);
  if (m.n_bulk > (1 << 20)) return fail(LENSIM_EINVAL, "n_bulk %d too large", m.n_bulk);
  m.n_defl = m.n_lens + m.n_bulk;
  m.n_prof = n_inline + m.n_bulk;
  for (int k = 0; k < n_inline; ++k) {
    const int t = model->types[k];
    m.types[k] = t;
    m.epl_index[k] = -1;
    const bool lens = k < m.n_lens;
    const bool lens_ok = t == LENSIM_EPL || t == LENSIM_SIE || t == LENSIM_NIE || t == LENSIM_NFW ||
                         t == LENSIM_TNFW || t == LENSIM_SHEAR || t == LENSIM_CONVERGENCE;
    const bool light_ok = t == LENSIM_SERSIC || t == LENSIM_SERSIC_ELLIPSE || t == LENSIM_GAUSSIAN;
    if (lens && !lens_ok) return fail(LENSIM_EUNSUPPORTED, "profile %d: type %d is not a supported deflector", k, t);
    if (!lens && !light_ok)
      return fail(LENSIM_EUNSUPPORTED, "profile %d: type %d is not a supported light profile", k, t);
    if (t == LENSIM_EPL) m.epl_index[k] = m.n_epl++;
    if (t == LENSIM_NFW || t == LENSIM_TNFW) m.has_halo = 1;
    if (lens && (t == LENSIM_SIE || t == LENSIM_NIE || t == LENSIM_NFW)) m.has_nie_nfw = 1;
  }
  if (m.n_bulk > 0) {
    m.bulk_type = model->bulk_type;
    m.bulk_types = model->d_bulk_types;
    if (!bulk_type_ok(m.bulk_type))
      return fail(LENSIM_EUNSUPPORTED, "bulk deflector type %d not supported (SIE, NIE, NFW, TNFW, SHEAR, CONVERGENCE)",
                  m.bulk_type);
    if (m.bulk_type == LENSIM_NULL && !m.bulk_types)
      return fail(LENSIM_EINVAL, "mixed bulk deflectors need d_bulk_types");
    if (m.bulk_type == LENSIM_NFW || m.bulk_type == LENSIM_TNFW || m.bulk_type == LENSIM_NULL) m.has_halo = 1;
    if (m.bulk_type == LENSIM_SIE || m.bulk_type == LENSIM_NIE || m.bulk_type == LENSIM_NFW ||
        m.bulk_type == LENSIM_NULL)
      m.has_nie_nfw = 1;
  }
  m.alpha_scaling = model->alpha_scaling;
  m.flux_scaling = model->flux_scaling;
  if (const lensim_multiplane *mp = model->multiplane) {
    if (mp->n_planes != m.n_defl + 1)
      return fail(LENSIM_EINVAL, "multiplane.n_planes %d must be the %d deflectors + 1", mp->n_planes, m.n_defl);
    if ((m.n_defl > 0 && !mp->d_order) || !mp->d_T_ij || !mp->d_T_z || !mp->d_r2p)
      return fail(LENSIM_EINVAL, "multiplane arrays are NULL");
    if (mp->stride != 0 && mp->stride < mp->n_planes) return fail(LENSIM_EINVAL, "multiplane stride too small");
    m.mp_n = mp->n_planes;
    m.mp_order = mp->d_order;
    m.mp_tij = mp->d_T_ij;
    m.mp_tz = mp->d_T_z;
    m.mp_r2p = mp->d_r2p;
    m.mp_stride = mp->stride;
    // with several planes a deflector's error moves the rays at every later
    // plane, where a strong lens magnifies it: the float form's render and
    // gradient kernels deflect the NFW family in double there too (the float
    // TNFW left logL 2e-6 off)
    if (m.has_halo) m.has_nie_nfw = 1;
  }
  return LENSIM_OK;
}

// ------------------------------------------------------------- kernels ---

__device__ inline bool bulk_type_ok_dev(int t) {
  return t == LENSIM_NULL || t == LENSIM_SIE || t == LENSIM_NIE || t == LENSIM_NFW || t == LENSIM_TNFW ||
         t == LENSIM_SHEAR || t == LENSIM_CONVERGENCE;
}

// Record (and EPL coefficients) number idx < n_img * n_prof, profile-major
// (idx = k * n_img + img), so the lanes of a wave share one profile type and
// take one branch of prepare_record; records are stored [img][k].
__device__ void prep_one(const ModelArgs &m, const double *__restrict__ params, int n_img, int idx,
                         double *__restrict__ recd, float *__restrict__ recf, double *__restrict__ coefd,
                         float *__restrict__ coeff, double *__restrict__ dcoefd, float *__restrict__ dcoeff) {
  const int k = idx / n_img, img = idx - k * n_img;
  const int ri = img * m.n_prof + k;
  if (k >= m.n_lens && k < m.n_defl && !m.bulk_type) CHECK_SPAN(m.bulk_types + (k - m.n_lens), sizeof(int32_t));
  int type = type_of(m, k);
  if (k >= m.n_lens && k < m.n_defl && !bulk_type_ok_dev(type)) type = LENSIM_NULL;
  const double *p = params + size_t(ri) * LENSIM_NPARAM;
  CHECK_SPAN(p, LENSIM_NPARAM * sizeof(double));
  double rec[kRec];
  double *coef = nullptr, *dcoef = nullptr;
  float *cf = nullptr, *dcf = nullptr;
  const int e = k < m.n_lens ? m.epl_index[k] : -1;
  if (e >= 0) {
    const size_t co = (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX;
    coef = coefd + co;
    cf = coeff + co;
    CHECK_SPAN(coef, LENSIM_EPL_NMAX * sizeof(double));
    CHECK_SPAN(cf, LENSIM_EPL_NMAX * sizeof(float));
    if (dcoefd) {
      dcoef = dcoefd + co;
      dcf = dcoeff + co;
      CHECK_SPAN(dcoef, LENSIM_EPL_NMAX * sizeof(double));
      CHECK_SPAN(dcf, LENSIM_EPL_NMAX * sizeof(float));
    }
  }
  prepare_record(type, p, rec, coef, m.sub, dcoef, cf, dcf);
  for (int i = 0; i < kRec; ++i) {
    AT(recd, size_t(ri) * kRec + i) = rec[i];
    AT(recf, size_t(ri) * kRec + i) = float(rec[i]);
  }
  if (type == LENSIM_EPL) {  // series lengths as int bits in the float copy
    AT(recf, size_t(ri) * kRec + EPL_NMAX) = __int_as_float(int(rec[EPL_NMAX]));
    AT(recf, size_t(ri) * kRec + EPL_NZ) = __int_as_float(int(rec[EPL_NZ]));
    AT(recf, size_t(ri) * kRec + EPL_NF) = __int_as_float(int(rec[EPL_NF]));
  }
}

__global__ void prep_kernel(ModelArgs m, const double *__restrict__ params, int n_img, double *__restrict__ recd,
                            float *__restrict__ recf, double *__restrict__ coefd, float *__restrict__ coeff,
                            double *__restrict__ dcoefd, float *__restrict__ dcoeff) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n_img * m.n_prof) prep_one(m, params, n_img, idx, recd, recf, coefd, coeff, dcoefd, dcoeff);
}

// Flip the kernel (so the convolution runs as a correlation) and store it
// transposed, column c of the flipped kernel contiguous: kt[c * kyp + a] =
// K[ky-1-a][kx-1-c], zero for a >= ky.
__device__ void psf_prep_one(const double *__restrict__ k, int kx, int ky, int kyp, double *__restrict__ out,
                             int flip, int idx) {
  const int c = idx / kyp, a = idx - c * kyp;
  if (!k) {  // identity kernel (no PSF / unconvolved)
    AT(out, idx) = (a == 0 && c == 0) ? 1.0 : 0.0;
    return;
  }
  // flip = 0 stores the kernel itself: the convolution then correlates with K,
  // which is the adjoint of convolving with K (zero padding)
  AT(out, idx) = (a < ky) ? (flip ? AT(k, (ky - 1 - a) * kx + (kx - 1 - c)) : AT(k, a * kx + c)) : 0.0;
}

__global__ void psf_prep_kernel(const double *__restrict__ k, int kx, int ky, int kyp, double *__restrict__ out,
                                int flip = 1) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < kyp * kx) psf_prep_one(k, kx, ky, kyp, out, flip, idx);
}

template <typename T>
struct RecPtrs {
  const double *recd;
  const T *rect;
  const double *coefd;
  const T *coeft;
};

template <typename T>
__device__ __forceinline__ RecPtrs<T> rec_ptrs(const ModelArgs &m, const double *recd, const float *recf,
                                               const double *coefd, const float *coeff, int img) {
  RecPtrs<T> r;
  r.recd = recd + size_t(img) * m.n_prof * kRec;
  CHECK_SPAN(r.recd, size_t(m.n_prof) * kRec * sizeof(double));
  if constexpr (sizeof(T) == sizeof(float)) {
    r.rect = recf + size_t(img) * m.n_prof * kRec;
    CHECK_SPAN(r.rect, size_t(m.n_prof) * kRec * sizeof(float));
    r.coeft = coeff;
  } else {
    r.rect = r.recd;
    r.coeft = coefd;
  }
  r.coefd = coefd;
  return r;
}

// The deflection of the bulk list (records n_lens .. n_defl - 1) added to (ax, ay).
template <typename T, bool HALO, bool DBL = false, bool DBL_TNFW = DBL>
__device__ __forceinline__ void deflect_bulk(const ModelArgs &m, const RecPtrs<T> &rp, double x, double y,
                                             double &ax, double &ay) {
  if (m.n_bulk == 0) return;
  if (m.bulk_type == LENSIM_TNFW) {
    // common case of substructure: one profile type, no per-deflector switch
    if constexpr (HALO) {
      for (int k = m.n_lens; k < m.n_defl; ++k) {
        if constexpr (DBL_TNFW) tnfw_defl<double>(Rec<double>{rp.recd + k * kRec, rp.recd + k * kRec}, x, y, ax, ay);
        else tnfw_defl<T>(Rec<T>{rp.recd + k * kRec, rp.rect + k * kRec}, x, y, ax, ay);
      }
    }
    return;
  }
  for (int k = m.n_lens; k < m.n_defl; ++k) {
    const int type = m.bulk_type ? m.bulk_type : AT(m.bulk_types, k - m.n_lens);
    const Rec<T> r{rp.recd + k * kRec, rp.rect + k * kRec};
    add_deflection<T, HALO, DBL, DBL_TNFW>(type == LENSIM_EPL ? LENSIM_NULL : type, r, nullptr, nullptr, x, y, ax, ay);
  }
}

template <typename T, bool HALO, bool DBL = false, bool DBL_TNFW = DBL>
__device__ __forceinline__ void deflect_all(const ModelArgs &m, const RecPtrs<T> &rp, int img, double x, double y,
                                            double &ax, double &ay) {
  for (int k = 0; k < m.n_lens; ++k) {
    const int type = m.types[k];
    const Rec<T> r{rp.recd + k * kRec, rp.rect + k * kRec};
    const int e = m.epl_index[k];
    const size_t co = (e >= 0) ? (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX : 0;
    if (e >= 0) CHECK_SPAN(rp.coeft + co, LENSIM_EPL_NMAX * sizeof(T));
    add_deflection<T, HALO, DBL, DBL_TNFW>(type, r, rp.coefd + co, rp.coeft + co, x, y, ax, ay);
  }
  deflect_bulk<T, HALO, DBL, DBL_TNFW>(m, rp, x, y, ax, ay);
}

// One deflector k (inline lens k < n_lens with its series coefficients, or bulk).
template <typename T, bool HALO, bool DBL, bool DBL_TNFW = DBL>
__device__ __forceinline__ void deflect_one(const ModelArgs &m, const RecPtrs<T> &rp, int img, int k, double x,
                                            double y, double &ax, double &ay) {
  const Rec<T> r{rp.recd + k * kRec, rp.rect + k * kRec};
  if (k < m.n_lens) {
    const int e = m.epl_index[k];
    const size_t co = (e >= 0) ? (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX : 0;
    if (e >= 0) CHECK_SPAN(rp.coeft + co, LENSIM_EPL_NMAX * sizeof(T));
    add_deflection<T, HALO, DBL, DBL_TNFW>(m.types[k], r, rp.coefd + co, rp.coeft + co, x, y, ax, ay);
  } else {
    const int type = m.bulk_type ? m.bulk_type : AT(m.bulk_types, k - m.n_lens);
    add_deflection<T, HALO, DBL, DBL_TNFW>(type == LENSIM_EPL ? LENSIM_NULL : type, r, nullptr, nullptr, x, y, ax, ay);
  }
}

// Source-plane position of the ray at angle (x, y): single plane
// beta = theta - alpha_scaling * alpha, or the multi-plane recursion of
// MultiPlaneBulk.ray_shooting (LensModel/MultiPlane/multi_plane_bulk.py:152-212).
template <typename T, bool HALO, bool DBL, bool DBL_TNFW = DBL>
__device__ __forceinline__ void trace_ray(const ModelArgs &m, const RecPtrs<T> &rp, int img, double x, double y,
                                          double &bx, double &by) {
  if (m.mp_n == 0) {
    double ax = 0.0, ay = 0.0;
    deflect_all<T, HALO, DBL, DBL_TNFW>(m, rp, img, x, y, ax, ay);
    bx = x - ax * m.alpha_scaling;
    by = y - ay * m.alpha_scaling;
    return;
  }
  const size_t o = size_t(img) * size_t(m.mp_stride);
  double angx = x, angy = y, px = 0.0, py = 0.0, tz = 1.0;
  for (int p = 0; p < m.mp_n; ++p) {
    const double dt = AT(m.mp_tij, o + p);
    tz = AT(m.mp_tz, o + p);
    px += angx * dt;
    py += angy * dt;
    if (p + 1 < m.mp_n) {
      double ax = 0.0, ay = 0.0;
      int k = AT(m.mp_order, p);
      k = k < 0 ? 0 : (k >= m.n_defl ? m.n_defl - 1 : k);  // a bad index cannot read outside the records
      deflect_one<T, HALO, DBL, DBL_TNFW>(m, rp, img, k, px / tz, py / tz, ax, ay);
      const double f = AT(m.mp_r2p, o + p);
      angx -= ax * f;
      angy -= ay * f;
    }
  }
  bx = px / tz;
  by = py / tz;
}

template <typename T>
__device__ __forceinline__ double light_sum(const ModelArgs &m, const RecPtrs<T> &rp, int k0, int k1, double x,
                                            double y) {
  double f = 0.0;
  for (int k = k0; k < k1; ++k) {
    const Rec<T> r{rp.recd + k * kRec, rp.rect + k * kRec};
    f += light_value<T>(m.types[k - m.n_bulk], r, x, y);
  }
  return f;
}

struct GridArgs {
  int nx, ny, s;
  double ra0, dec0, m00, m01, m10, m11;
  double scale;  // pixel_width^2 / s^2
  const uint8_t *mask;  // evaluation mask of the image pixels (NULL: all), [mask_ny][mask_nx]
  int mask_nx, mask_s;  // pixel (i, j) of this grid is mask pixel (i / mask_s, j / mask_s)
  // formed once on the host (grid_consts), so no thread divides: 1/s, the
  // first sub-ray's offset -1/2 + 1/(2 s) (make_subgrid, ImSim/Numerics/grid.py:60-62),
  // 1/nx and 1/s^2
  double inv_s, o0, inv_nx, inv_ss;
};

void grid_consts(GridArgs &g) {
  g.inv_s = 1.0 / g.s;
  g.o0 = -0.5 + 1.0 / (2.0 * g.s);
  g.inv_nx = 1.0 / g.nx;
  g.inv_ss = 1.0 / (double(g.s) * double(g.s));
}

// q = n / d and n - q d for 0 <= n < 2^30 and 1 <= d, from inv_d = 1/d in
// double: (n + 1/2) / d lies at least 1/(2d) from an integer, and the
// product's rounding error is below n 2^-52 < 2^-22 / d, so truncation gives q.
__device__ __forceinline__ int div_by(int n, int d, double inv_d, int &rem) {
  const int q = int((double(n) + 0.5) * inv_d);
  rem = n - q * d;
  return q;
}

// Whether grid pixel (i, j) is evaluated (RegularGrid flux_evaluate_indexes,
// ImSim/Numerics/grid.py:47-58: the rays of other pixels are not shot and
// their flux is zero).
__device__ __forceinline__ bool evaluated(const GridArgs &g, int i, int j) {
  return !g.mask || AT(g.mask, size_t(j / g.mask_s) * g.mask_nx + i / g.mask_s) != 0;
}

// The maximum over the 64 lanes of a wave (every lane active) of a
// non-negative float, by the integer max of the bit patterns (ordered as the
// values are): DPP quad permutations and row rotations give each row of 16
// lanes its maximum in every lane, readlane the four rows'.  Eight VALU
// instructions where a shuffle butterfly takes about fifty.
__device__ __forceinline__ float wave_max_nonneg(float x) {
  int v = __float_as_int(x);
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1, 0, 3, 2]
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2, 3, 0, 1]
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x124, 0xF, 0xF, false));  // row_ror:4
  v = max(v, __builtin_amdgcn_update_dpp(0, v, 0x128, 0xF, 0xF, false));  // row_ror:8
  const int r = max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
                    max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
  return __int_as_float(r);
}

// The block's maximum of the threads' non-negative x (every thread of the
// block reaches this call) into *slot: one slot per block, no atomics.  The
// MFMA convolution scales its fp16 operands by the image maximum.
__device__ __forceinline__ void block_max_store(float x, float *slot) {
  __shared__ float wmax[256 / 64];
  const float mv = wave_max_nonneg(x);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = mv;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = wmax[0];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) b = fmaxf(b, wmax[w]);
    AT(slot, 0) = b;
  }
}

template <typename T, bool HALO, bool DBL>
__global__ __launch_bounds__(256) void render_kernel(ModelArgs m, GridArgs g, const double *__restrict__ recd,
                                                     const float *__restrict__ recf,
                                                     const double *__restrict__ coefd,
                                                     const float *__restrict__ coeff, unsigned flags,
                                                     T *__restrict__ out, float *__restrict__ blkmax) {
  const int img = blockIdx.y;
  const int npix = g.nx * g.ny;
  const int pix0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = pix0 < npix;
  const int pix = active ? pix0 : npix - 1;  // inactive lanes stay for the wave reduction
  int i;
  const int j = div_by(pix, g.nx, g.inv_nx, i);
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  // pixel centre (lenstronomy grid_from_coordinate_transform order)
  const double ra_c = i * g.m00 + j * g.m01 + g.ra0;
  const double dec_c = i * g.m10 + j * g.m11 + g.dec0;
  const double inv_s = g.inv_s;
  const double o0 = g.o0;
  const bool do_src = (flags & LENSIM_IMAGE_SOURCE) && m.n_source > 0;
  const bool do_light = (flags & LENSIM_IMAGE_LENS_LIGHT) && m.n_light > 0;
  const int k_src0 = m.n_defl, k_src1 = m.n_defl + m.n_source, k_l1 = m.n_prof;
  double acc = 0.0;
  const bool eval = active && evaluated(g, i, j);
  for (int sy = 0; eval && sy < g.s; ++sy) {
    const double oy = o0 + sy * inv_s;
    for (int sx = 0; sx < g.s; ++sx) {
      const double ox = o0 + sx * inv_s;
      // make_subgrid offsets (ImSim/Numerics/grid.py:60-62)
      const double x = ra_c + g.m00 * ox + g.m01 * oy;
      const double y = dec_c + g.m10 * ox + g.m11 * oy;
      if (do_src) {
        double bx, by;
        trace_ray<T, HALO, DBL>(m, rp, img, x, y, bx, by);
        acc += light_sum<T>(m, rp, k_src0, k_src1, bx, by);
      }
      if (do_light) acc += light_sum<T>(m, rp, k_src1, k_l1, x, y);
    }
  }
  const T v = T(acc * (g.scale * m.flux_scaling));
  if (active) AT(out, size_t(img) * npix + pix) = v;
  if (blkmax) block_max_store(fabsf(float(v)), blkmax + size_t(img) * gridDim.x + blockIdx.x);
}

// The coordinates of a compile-time deflector that the sub-ray loop advances:
// Z = (q xr, yr) for the EPL, (xr, yr) for SIE / NIE (major-axis frame), the
// offset (dx, dy) from the centre for NFW.  They are affine in the ray
// position, so after their value at the pixel's first sub-ray each sub-ray
// adds a step (xs, ys) and each row a row step (xr_, yr_).
template <int TYPE>
struct DeflCoords {
  double a, b, sa, sb, ra, rb;
  __device__ __forceinline__ void init(const double *rd, const SubSteps &st, double x0, double y0) {
    const double dx = x0 - rd[R_CX], dy = y0 - rd[R_CY];
    if constexpr (TYPE == LENSIM_EPL) {
      const double c = rd[EPL_C], s = rd[EPL_S];
      a = rd[EPL_Q] * (dx * c + dy * s);
      b = -dx * s + dy * c;
      sa = rd[EPL_ZSX];
      sb = rd[EPL_ZSXI];
      ra = rd[EPL_ZRR];
      rb = rd[EPL_ZRRI];
    } else if constexpr (TYPE == LENSIM_NIE) {
      const double c = rd[NIE_C], s = rd[NIE_S];
      a = dx * c + dy * s;
      b = -dx * s + dy * c;
      sa = rd[NIE_XSX];
      sb = rd[NIE_YSX];
      ra = rd[NIE_XRR];
      rb = rd[NIE_YRR];
    } else {
      a = dx;
      b = dy;
      sa = st.sx[0];
      sb = st.sx[1];
      ra = st.rr[0];
      rb = st.rr[1];
    }
  }
  __device__ __forceinline__ void step() {
    a += sa;
    b += sb;
  }
  __device__ __forceinline__ void row() {
    a += ra;
    b += rb;
  }
};

// The exact position of sub-ray (sx, sy) of the pixel centred at (ra_c, dec_c).
struct SubRay {
  double ra_c, dec_c, o0, inv_s;
  int sx, sy;
  __device__ __forceinline__ void pos(const GridArgs &g, double &x, double &y) const {
    // the counters pass through readfirstlane (they are uniform), which keeps
    // the compiler from carrying them as float64 induction variables for
    // this rarely taken branch
    const double ox = o0 + __builtin_amdgcn_readfirstlane(sx) * inv_s;
    const double oy = o0 + __builtin_amdgcn_readfirstlane(sy) * inv_s;
    x = ra_c + g.m00 * ox + g.m01 * oy;
    y = dec_c + g.m10 * ox + g.m11 * oy;
  }
};

// beta -= alpha_scaling alpha of one compile-time deflector at its advanced
// coordinates; the EPL and SIE / NIE rotate back with (alpha_scaling cos phi,
// alpha_scaling sin phi) folded into the update.  An EPL sub-ray within 1e-12
// of the centre recomputes Z from its exact position (so that a ray exactly
// at the centre gets the reference's zero deflection, not the direction of
// the advanced coordinates' rounding).
template <typename T, int TYPE>
__device__ __forceinline__ void defl_fast(const Rec<T> &r, const T *coef_t, const double *coef_d,
                                          DeflCoords<TYPE> &u, double asc, const GridArgs &g, const SubRay &sr,
                                          double &bx, double &by) {
  if constexpr (TYPE == LENSIM_EPL) {
    double R2 = fma(u.a, u.a, u.b * u.b);
    if (R2 < 1e-24) {  // exact Z (the advance continues from it)
      double x, y;
      sr.pos(g, x, y);
      const double dx = x - r.d[R_CX], dy = y - r.d[R_CY];
      const double c = r.d[EPL_C], s = r.d[EPL_S];
      u.a = r.d[EPL_Q] * (dx * c + dy * s);
      u.b = -dx * s + dy * c;
      R2 = fma(u.a, u.a, u.b * u.b);
    }
    double fr, fi;
    if constexpr (sizeof(T) == sizeof(float))
      epl_major_f32(r, coef_t, coef_d, u.a, u.b, R2, fr, fi);
    else
      epl_major_f64(r, coef_d, u.a, u.b, R2, fr, fi);
    bx = fma(-r.d[EPL_AC], fr, fma(r.d[EPL_AS], fi, bx));
    by = fma(-r.d[EPL_AS], fr, fma(-r.d[EPL_AC], fi, by));
  } else if constexpr (TYPE == LENSIM_NIE) {
    double fx, fy;
    nie_major_dbl(r.d, u.a, u.b, fx, fy);
    bx = fma(-r.d[NIE_AC], fx, fma(r.d[NIE_AS], fy, bx));
    by = fma(-r.d[NIE_AS], fx, fma(-r.d[NIE_AC], fy, by));
  } else if constexpr (TYPE == LENSIM_NFW) {
    const double f = asc * nfw_factor_dbl(r.d, u.a, u.b);
    bx = fma(-f, u.a, bx);
    by = fma(-f, u.b, by);
  }
}

// The render for the common model families, the profile list fixed at compile
// time: deflector D0 (EPL, SIE / NIE or NFW) and optionally D1 (SHEAR, SIE /
// NIE or NFW; 0: none), in that order, then NSRC Sersic sources and NLL Sersic
// lens lights; one plane, no bulk list.  Per profile it is render_kernel's
// arithmetic, rearranged so that everything affine in the ray position is
// advanced by one addition per sub-ray instead of formed from the position:
// the deflectors' coordinates (DeflCoords), the lens lights' scaled
// elliptical coordinates, and w = theta - alpha_scaling alpha_SHEAR(theta)
// (the external shear is linear, so beta = w - alpha_scaling (alpha of the
// other deflectors)).  The records hold the steps (prepare_record's
// SubSteps).  In the float form each Sersic profile's values of a pixel are
// summed in float and multiplied by its amplitude in double at the end.
template <typename T, int D0, int D1, int NSRC, int NLL>
__global__ __launch_bounds__(256) void render_fast_kernel(ModelArgs m, GridArgs g, const double *__restrict__ recd,
                                                          const float *__restrict__ recf,
                                                          const double *__restrict__ coefd,
                                                          const float *__restrict__ coeff, unsigned flags,
                                                          T *__restrict__ out, float *__restrict__ blkmax) {
  constexpr bool kF32 = sizeof(T) == sizeof(float);
  static_assert(NSRC == 1, "the compile-time render takes one source");
  const int img = blockIdx.y;
  const int npix = g.nx * g.ny;
  const int pix0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = pix0 < npix;
  const int pix = active ? pix0 : npix - 1;  // inactive lanes stay for the wave reduction
  int i;
  const int j = div_by(pix, g.nx, g.inv_nx, i);
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  constexpr int KS = D1 ? 2 : 1;  // first source record
  const Rec<T> r0{rp.recd, rp.rect};
  const Rec<T> r1{rp.recd + kRec, rp.rect + kRec};
  const size_t co = size_t(img) * m.n_epl * LENSIM_EPL_NMAX;  // D0's coefficients when it is the EPL (epl_index 0)
  // the pixel's first sub-ray (make_subgrid offsets, ImSim/Numerics/grid.py:60-62)
  const double ra_c = i * g.m00 + j * g.m01 + g.ra0;
  const double dec_c = i * g.m10 + j * g.m11 + g.dec0;
  const double inv_s = g.inv_s;
  const double o0 = g.o0;
  const double x0 = ra_c + g.m00 * o0 + g.m01 * o0;
  const double y0 = dec_c + g.m10 * o0 + g.m11 * o0;
  const bool do_src = (flags & LENSIM_IMAGE_SOURCE) && NSRC > 0;
  const bool do_light = (flags & LENSIM_IMAGE_LENS_LIGHT) && NLL > 0;
  const double asc = m.alpha_scaling;
  // w = theta - asc alpha_SHEAR(theta) and its steps
  double wx, wy, wsx, wsy, wrx, wry;
  if constexpr (D1 == LENSIM_SHEAR) {
    const double dx = x0 - r1.d[R_CX], dy = y0 - r1.d[R_CY];
    const double g1 = r1.d[SH_G1], g2 = r1.d[SH_G2];
    wx = x0 - asc * (g1 * dx + g2 * dy);
    wy = y0 - asc * (g2 * dx - g1 * dy);
    wsx = r1.d[SH_WSX];
    wsy = r1.d[SH_WSY];
    wrx = r1.d[SH_WRX];
    wry = r1.d[SH_WRY];
  } else {
    wx = x0;
    wy = y0;
    wsx = m.sub.sx[0];
    wsy = m.sub.sx[1];
    wrx = m.sub.rr[0];
    wry = m.sub.rr[1];
  }
  // relative to the source's centre, so beta - c_source comes out directly
  const double *rds = rp.recd + KS * kRec;
  wx -= rds[R_CX];
  wy -= rds[R_CY];
  // the Sersic floors: canonicalised once for the FP64 form's fmax, as float
  // bits for the float form's integer max
  const double umin_s = __builtin_canonicalize(rds[SE_UMIN]);
  const int uminb_s = __float_as_int(float(rds[SE_UMIN]));
  double umin_l[NLL > 0 ? NLL : 1];
  int uminb_l[NLL > 0 ? NLL : 1];
#pragma unroll
  for (int k = 0; k < NLL; ++k) {
    umin_l[k] = __builtin_canonicalize(rp.recd[(KS + NSRC + k) * kRec + SE_UMIN]);
    uminb_l[k] = __float_as_int(float(rp.recd[(KS + NSRC + k) * kRec + SE_UMIN]));
  }
  DeflCoords<D0> u0;
  u0.init(r0.d, m.sub, x0, y0);
  constexpr int D1C = (D1 == LENSIM_SHEAR || D1 == 0) ? LENSIM_NFW : D1;  // unused when D1 is SHEAR or none
  DeflCoords<D1C> u1;
  if constexpr (D1 == LENSIM_NIE || D1 == LENSIM_NFW) u1.init(r1.d, m.sub, x0, y0);
  // the lens lights' scaled elliptical coordinates
  double lx[NLL > 0 ? NLL : 1], ly[NLL > 0 ? NLL : 1];
#pragma unroll
  for (int k = 0; k < NLL; ++k) {
    const double *rd = rp.recd + (KS + NSRC + k) * kRec;
    const double dx = x0 - rd[R_CX], dy = y0 - rd[R_CY];
    lx[k] = fma(rd[SE_M00S], dx, rd[SE_M01S] * dy);
    ly[k] = fma(rd[SE_M01S], dx, rd[SE_M11S] * dy);
  }
  // per-profile sums of the Sersic values (float form) or the flux (FP64 form)
  float fs[NSRC > 0 ? NSRC : 1], fl[NLL > 0 ? NLL : 1];
#pragma unroll
  for (int k = 0; k < NSRC; ++k) fs[k] = 0.f;
#pragma unroll
  for (int k = 0; k < NLL; ++k) fl[k] = 0.f;
  double acc = 0.0;
  const bool eval = active && evaluated(g, i, j);
  for (int sy = 0; eval && sy < g.s; ++sy) {
    for (int sx = 0; sx < g.s; ++sx) {
      if (do_src) {
        const SubRay sr{ra_c, dec_c, o0, inv_s, sx, sy};
        double bx = wx, by = wy;
        defl_fast<T, D0>(r0, rp.coeft + co, rp.coefd + co, u0, asc, g, sr, bx, by);
        if constexpr (D1 == LENSIM_NIE || D1 == LENSIM_NFW)
          defl_fast<T, D1C>(r1, nullptr, nullptr, u1, asc, g, sr, bx, by);
        // (bx, by) = beta - c_source
        const double xs = fma(rds[SE_M00S], bx, rds[SE_M01S] * by);
        const double ys = fma(rds[SE_M01S], bx, rds[SE_M11S] * by);
        if constexpr (kF32)
          fs[0] += sersic_u_f32(rds, reinterpret_cast<const float *>(rp.rect) + KS * kRec, xs, ys, uminb_s);
        else
          acc += rds[SE_AMP] * sersic_u_f64(rds, xs, ys, umin_s);
      }
      if (do_light) {
#pragma unroll
        for (int k = 0; k < NLL; ++k) {
          const double *rd = rp.recd + (KS + NSRC + k) * kRec;
          if constexpr (kF32)
            fl[k] += sersic_u_f32(rd, reinterpret_cast<const float *>(rp.rect) + (KS + NSRC + k) * kRec, lx[k],
                                  ly[k], uminb_l[k]);
          else
            acc += rd[SE_AMP] * sersic_u_f64(rd, lx[k], ly[k], umin_l[k]);
        }
      }
      // next sub-ray
      wx += wsx;
      wy += wsy;
      u0.step();
      if constexpr (D1 == LENSIM_NIE || D1 == LENSIM_NFW) u1.step();
#pragma unroll
      for (int k = 0; k < NLL; ++k) {
        const double *rd = rp.recd + (KS + NSRC + k) * kRec;
        lx[k] += rd[SE_XSX];
        ly[k] += rd[SE_YSX];
      }
    }
    // next row
    wx += wrx;
    wy += wry;
    u0.row();
    if constexpr (D1 == LENSIM_NIE || D1 == LENSIM_NFW) u1.row();
#pragma unroll
    for (int k = 0; k < NLL; ++k) {
      const double *rd = rp.recd + (KS + NSRC + k) * kRec;
      lx[k] += rd[SE_XRR];
      ly[k] += rd[SE_YRR];
    }
  }
  if constexpr (kF32) {
#pragma unroll
    for (int k = 0; k < NSRC; ++k) acc = fma(rp.recd[(KS + k) * kRec + SE_AMP], double(fs[k]), acc);
#pragma unroll
    for (int k = 0; k < NLL; ++k) acc = fma(rp.recd[(KS + NSRC + k) * kRec + SE_AMP], double(fl[k]), acc);
  }
  const T v = T(acc * (g.scale * m.flux_scaling));
  if (active) AT(out, size_t(img) * npix + pix) = v;
  if (blkmax) block_max_store(fabsf(float(v)), blkmax + size_t(img) * gridDim.x + blockIdx.x);
}

int check_launch(const char *where);

// The per-256-pixel block maxima of |image| that render_kernel stores for the
// MFMA convolution, from an image rendered by a kernel that keeps none.
template <typename T>
__global__ __launch_bounds__(256) void block_max_kernel(const T *__restrict__ v, int npix,
                                                        float *__restrict__ blkmax) {
  const int img = blockIdx.y;
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  const float a = pix < npix ? fabsf(float(AT(v, size_t(img) * npix + pix))) : 0.f;
  block_max_store(a, blkmax + size_t(img) * gridDim.x + blockIdx.x);
}

// Lanes per ray of the multi-plane kernels that share a plane's deflectors
// among a quad (raytrace_mp_kernel, render_mp_kernel).
constexpr int kMpG = 4;

template <typename T, bool HALO, bool DBL>
__global__ __launch_bounds__(256) void render_mp_kernel(ModelArgs m, GridArgs g, const double *__restrict__ recd,
                                                        const float *__restrict__ recf,
                                                        const double *__restrict__ coefd,
                                                        const float *__restrict__ coeff, unsigned flags,
                                                        T *__restrict__ out);

// Launches render_fast_kernel when the model is one of its compile-time lists
// (EPL, SIE / NIE or NFW, each with or without SHEAR after it, NFW + SIE / NIE
// in either order; one Sersic source, none or one Sersic lens light), else
// render_kernel.
template <typename T>
int launch_render(const ModelArgs &m, const GridArgs &g, const double *recd, const float *recf, const double *coefd,
                  const float *coeff, unsigned flags, T *out, float *blkmax, int n_img, hipStream_t st) {
  const dim3 rgrid((g.nx * g.ny + 255) / 256, n_img);
  auto sersic = [](int t) { return t == LENSIM_SERSIC || t == LENSIM_SERSIC_ELLIPSE; };
  auto nie = [](int t) { return t == LENSIM_SIE ? LENSIM_NIE : t; };
  bool fast = m.mp_n == 0 && m.n_bulk == 0 && (m.n_lens == 1 || m.n_lens == 2) && m.n_source == 1 &&
              (m.n_light == 0 || m.n_light == 1);
  int d0 = 0, d1 = 0;
  if (fast) {
    d0 = nie(m.types[0]);
    d1 = m.n_lens == 2 ? nie(m.types[1]) : 0;
    const bool pair = ((d0 == LENSIM_EPL || d0 == LENSIM_NIE || d0 == LENSIM_NFW) && (d1 == 0 || d1 == LENSIM_SHEAR)) ||
                      (d0 == LENSIM_NFW && d1 == LENSIM_NIE) || (d0 == LENSIM_NIE && d1 == LENSIM_NFW);
    fast = pair && (d0 != LENSIM_EPL || m.epl_index[0] == 0);
    for (int k = m.n_lens; fast && k < m.n_lens + m.n_source + m.n_light; ++k) fast = sersic(m.types[k]);
  }
  // a multi-plane launch of fewer than 2 waves per SIMD at one lane per pixel
  // takes a quad per pixel (render_mp_kernel; the block maxima for the MFMA
  // convolution follow from the image in a second launch)
  const bool quad = m.mp_n && size_t(g.nx) * g.ny * size_t(n_img) < 131072;
  if (!fast && quad) {
    const dim3 qgrid((unsigned)((size_t(g.nx) * g.ny * kMpG + 255) / 256), n_img);
#define LENSIM_RENDER(H, D) \
  render_mp_kernel<T, H, D><<<qgrid, 256, 0, st>>>(m, g, recd, recf, coefd, coeff, flags, out)
    bool done = false;
    if constexpr (sizeof(T) == sizeof(float)) {
      if (m.has_nie_nfw) {
        if (m.has_halo) LENSIM_RENDER(true, true);
        else LENSIM_RENDER(false, true);
        done = true;
      }
    }
    if (!done) {
      if (m.has_halo) LENSIM_RENDER(true, false);
      else LENSIM_RENDER(false, false);
    }
#undef LENSIM_RENDER
    int rc = check_launch("render_mp_kernel");
    if (rc || !blkmax) return rc;
    block_max_kernel<T><<<rgrid, 256, 0, st>>>(out, g.nx * g.ny, blkmax);
    return check_launch("block_max_kernel");
  }
  if (!fast) {
    // SIE / NIE and NFW deflect in double in the float form too; the kernel
    // is instantiated both ways so other models keep the float forms' registers
#define LENSIM_RENDER(H, D) \
  render_kernel<T, H, D><<<rgrid, 256, 0, st>>>(m, g, recd, recf, coefd, coeff, flags, out, blkmax)
    bool done = false;
    if constexpr (sizeof(T) == sizeof(float)) {
      if (m.has_nie_nfw) {
        if (m.has_halo) LENSIM_RENDER(true, true);
        else LENSIM_RENDER(false, true);
        done = true;
      }
    }
    if (!done) {
      if (m.has_halo) LENSIM_RENDER(true, false);
      else LENSIM_RENDER(false, false);
    }
#undef LENSIM_RENDER
    return check_launch("render_kernel");
  }
  const bool ll = m.n_light == 1;
#define LENSIM_FAST(A, B)                                                                                       \
  do {                                                                                                          \
    if (ll)                                                                                                     \
      render_fast_kernel<T, A, B, 1, 1><<<rgrid, 256, 0, st>>>(m, g, recd, recf, coefd, coeff, flags, out, blkmax); \
    else                                                                                                        \
      render_fast_kernel<T, A, B, 1, 0><<<rgrid, 256, 0, st>>>(m, g, recd, recf, coefd, coeff, flags, out, blkmax); \
  } while (0)
  if (d0 == LENSIM_EPL && d1 == LENSIM_SHEAR) LENSIM_FAST(LENSIM_EPL, LENSIM_SHEAR);
  else if (d0 == LENSIM_EPL) LENSIM_FAST(LENSIM_EPL, 0);
  else if (d0 == LENSIM_NIE && d1 == LENSIM_SHEAR) LENSIM_FAST(LENSIM_NIE, LENSIM_SHEAR);
  else if (d0 == LENSIM_NIE && d1 == LENSIM_NFW) LENSIM_FAST(LENSIM_NIE, LENSIM_NFW);
  else if (d0 == LENSIM_NIE) LENSIM_FAST(LENSIM_NIE, 0);
  else if (d0 == LENSIM_NFW && d1 == LENSIM_SHEAR) LENSIM_FAST(LENSIM_NFW, LENSIM_SHEAR);
  else if (d0 == LENSIM_NFW && d1 == 0) LENSIM_FAST(LENSIM_NFW, 0);
  else LENSIM_FAST(LENSIM_NFW, LENSIM_NIE);
#undef LENSIM_FAST
  return check_launch("render_fast_kernel");
}

template <typename T, bool HALO, bool DBL>
__global__ __launch_bounds__(256) void raytrace_kernel(ModelArgs m, const double *__restrict__ recd,
                                                       const float *__restrict__ recf,
                                                       const double *__restrict__ coefd,
                                                       const float *__restrict__ coeff, const double *__restrict__ xs,
                                                       const double *__restrict__ ys, int64_t n_ray,
                                                       double *__restrict__ bx, double *__restrict__ by) {
  const int img = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_ray) return;
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  double tx, ty;
  trace_ray<T, HALO, DBL, false>(m, rp, img, AT(xs, i), AT(ys, i), tx, ty);
  AT(bx, size_t(img) * n_ray + i) = tx;
  AT(by, size_t(img) * n_ray + i) = ty;
}

// A double moved across the lanes of each quad by a DPP quad permutation.
template <int CTRL>
__device__ __forceinline__ double dpp_quad_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// The sum over the 4 lanes of a quad (every lane active), (v0 + v1) + (v2 + v3)
// in every lane: the additions commute exactly, so the four lanes hold the
// same bits.
__device__ __forceinline__ double quad_sum(double v) {
  v += dpp_quad_f64<0xB1>(v);  // quad_perm [1, 0, 3, 2]
  v += dpp_quad_f64<0x4E>(v);  // quad_perm [2, 3, 0, 1]
  return v;
}

// The sum over a wave (every lane active), uniform: the quads' sums, then
// the row's (row_ror 4 and 8 within each row of 16 lanes), then the four rows'
// sums read from their first lanes and added in order.
__device__ __forceinline__ double wave_sum(double v) {
  v = quad_sum(v);
  v += dpp_quad_f64<0x124>(v);  // row_ror:4
  v += dpp_quad_f64<0x128>(v);  // row_ror:8
  auto row = [&](int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
  };
  return (row(0) + row(16)) + (row(32) + row(48));
}

// Multi-plane ray tracing (MultiPlaneBulk.ray_shooting,
// LensModel/MultiPlane/multi_plane_bulk.py:152-212) with kMpG lanes per ray.
// Deflectors on one plane follow each other in the plane order with T_ij = 0
// between them, so the ray's comoving position does not move across them and
// they all deflect it at one position: lane j of the ray's quad evaluates
// deflector p + j of such a run (up to kMpG at a time) and a quad sum adds the
// r2p-weighted deflections, which the reference subtracts one after another.
// The order and the T arrays are the same for every ray of an image, so the
// runs are uniform across the wave (scalar loads and branches).  A ray's
// solitary plane-by-plane chain of record loads and deflections is kMpG times
// shorter, and an image of 400^2 rays fills 10 waves per SIMD instead of 2.5.
// The float form takes it (the line-of-sight workload: 9.1 -> 8.5 ms); the
// FP64 form's software transcendentals keep the vector ALU busy at 2.5 waves
// per SIMD already, and the gathers and quad sums made it slower (10.8 ->
// 11.9 ms), so it keeps one lane per ray.

template <typename T, bool HALO, bool DBL>
__global__ __launch_bounds__(256) void raytrace_mp_kernel(ModelArgs m, const double *__restrict__ recd,
                                                          const float *__restrict__ recf,
                                                          const double *__restrict__ coefd,
                                                          const float *__restrict__ coeff,
                                                          const double *__restrict__ xs,
                                                          const double *__restrict__ ys, int64_t n_ray,
                                                          double *__restrict__ bx, double *__restrict__ by) {
  const int img = blockIdx.y;
  const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const int slot = threadIdx.x & (kMpG - 1);
  const int64_t ray0 = t / kMpG;
  const bool active = ray0 < n_ray;
  const int64_t ray = active ? ray0 : n_ray - 1;  // a quad past the end follows the last ray (for the sums)
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  const size_t o = size_t(img) * size_t(m.mp_stride);
  const double *tij = m.mp_tij + o, *tzs = m.mp_tz + o, *r2p = m.mp_r2p + o;
  double angx = AT(xs, ray), angy = AT(ys, ray), px = 0.0, py = 0.0;
  const int nd = m.mp_n - 1;  // planes with a deflector; the last one is the source's
  int p = 0;
  while (p < nd) {
    const double dt = AT(tij, p);
    px += angx * dt;
    py += angy * dt;
    // deflectors p .. p + mc - 1 see the same position
    int mc = 1;
    while (mc < kMpG && p + mc < nd && AT(tij, p + mc) == 0.0) ++mc;
    double ax = 0.0, ay = 0.0;
    if (slot < mc) {
      const int q = p + slot;
      int k = AT(m.mp_order, q);
      k = k < 0 ? 0 : (k >= m.n_defl ? m.n_defl - 1 : k);  // a bad index cannot read outside the records
      const double tz = AT(tzs, q);
      double dx = 0.0, dy = 0.0;
      deflect_one<T, HALO, DBL, false>(m, rp, img, k, px / tz, py / tz, dx, dy);
      const double f = AT(r2p, q);
      ax = dx * f;
      ay = dy * f;
    }
    angx -= quad_sum(ax);
    angy -= quad_sum(ay);
    p += mc;
  }
  const double dt = AT(tij, nd);
  px += angx * dt;
  py += angy * dt;
  const double tz = AT(tzs, nd);
  if (active && slot == 0) {
    AT(bx, size_t(img) * n_ray + ray) = px / tz;
    AT(by, size_t(img) * n_ray + ray) = py / tz;
  }
}

// render_kernel for multi-plane lists with a quad per pixel: the quad walks
// the pixel's sub-rays together and shares each plane's deflectors among its
// lanes as raytrace_mp_kernel does (with render_kernel's deflection
// arithmetic, trace_ray's), so every lane ends with the sub-ray's beta; every
// lane adds the light, lane 0 stores.  For launches too small to fill the
// SIMDs at one lane per pixel (a gradient's image of one line-of-sight scene).
template <typename T, bool HALO, bool DBL>
__global__ __launch_bounds__(256) void render_mp_kernel(ModelArgs m, GridArgs g, const double *__restrict__ recd,
                                                        const float *__restrict__ recf,
                                                        const double *__restrict__ coefd,
                                                        const float *__restrict__ coeff, unsigned flags,
                                                        T *__restrict__ out) {
  const int img = blockIdx.y;
  const int npix = g.nx * g.ny;
  const int slot = threadIdx.x & (kMpG - 1);
  const int pix0 = int((size_t(blockIdx.x) * blockDim.x + threadIdx.x) / kMpG);
  const bool active = pix0 < npix;
  const int pix = active ? pix0 : npix - 1;
  int i;
  const int j = div_by(pix, g.nx, g.inv_nx, i);
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  const double ra_c = i * g.m00 + j * g.m01 + g.ra0;
  const double dec_c = i * g.m10 + j * g.m11 + g.dec0;
  const double inv_s = g.inv_s;
  const double o0 = g.o0;
  const bool do_src = (flags & LENSIM_IMAGE_SOURCE) && m.n_source > 0;
  const bool do_light = (flags & LENSIM_IMAGE_LENS_LIGHT) && m.n_light > 0;
  const int k_src0 = m.n_defl, k_src1 = m.n_defl + m.n_source, k_l1 = m.n_prof;
  const size_t o = size_t(img) * size_t(m.mp_stride);
  const double *tij = m.mp_tij + o, *tzs = m.mp_tz + o, *r2p = m.mp_r2p + o;
  const int nd = m.mp_n - 1;  // planes with a deflector; the last one is the source's
  double acc = 0.0;
  const bool eval = active && evaluated(g, i, j);  // one value per quad (one pixel)
  for (int sy = 0; eval && sy < g.s; ++sy) {
    const double oy = o0 + sy * inv_s;
    for (int sx = 0; sx < g.s; ++sx) {
      const double ox = o0 + sx * inv_s;
      const double x = ra_c + g.m00 * ox + g.m01 * oy;
      const double y = dec_c + g.m10 * ox + g.m11 * oy;
      if (do_src) {
        double angx = x, angy = y, px = 0.0, py = 0.0;
        int p = 0;
        while (p < nd) {
          const double dt = AT(tij, p);
          px += angx * dt;
          py += angy * dt;
          int mc = 1;  // deflectors p .. p + mc - 1 see the same position
          while (mc < kMpG && p + mc < nd && AT(tij, p + mc) == 0.0) ++mc;
          double ax = 0.0, ay = 0.0;
          if (slot < mc) {
            const int q = p + slot;
            int k = AT(m.mp_order, q);
            k = k < 0 ? 0 : (k >= m.n_defl ? m.n_defl - 1 : k);  // a bad index cannot read outside the records
            const double tz = AT(tzs, q);
            double dx = 0.0, dy = 0.0;
            deflect_one<T, HALO, DBL, DBL>(m, rp, img, k, px / tz, py / tz, dx, dy);
            const double f = AT(r2p, q);
            ax = dx * f;
            ay = dy * f;
          }
          angx -= quad_sum(ax);
          angy -= quad_sum(ay);
          p += mc;
        }
        const double dt = AT(tij, nd);
        px += angx * dt;
        py += angy * dt;
        const double tz = AT(tzs, nd);
        acc += light_sum<T>(m, rp, k_src0, k_src1, px / tz, py / tz);
      }
      if (do_light) acc += light_sum<T>(m, rp, k_src1, k_l1, x, y);
    }
  }
  const T v = T(acc * (g.scale * m.flux_scaling));
  if (active && slot == 0) AT(out, size_t(img) * npix + pix) = v;
}

template <typename T>
__global__ __launch_bounds__(256) void light_kernel(ModelArgs m, const double *__restrict__ recd,
                                                    const float *__restrict__ recf, int k0, int k1,
                                                    const double *__restrict__ xs, const double *__restrict__ ys,
                                                    int64_t n_ray, double *__restrict__ flux) {
  const int img = blockIdx.y;
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_ray) return;
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, nullptr, nullptr, img);
  AT(flux, size_t(img) * n_ray + i) = light_sum<T>(m, rp, k0, k1, AT(xs, i), AT(ys, i));
}

struct ConvArgs {
  int nx, ny, kx, ky, kyp, edge, n_tiles_x;  // kyp: kernel height padded to kWin
  const double *data, *noise, *expo, *mask;
  double exposure_time, bkg2;
  double inv_t;  // 1 / exposure_time, formed on the host
  void *gout;  // optional dlogL/dmodel per pixel (compute type)
  int64_t dstride;  // data / noise elements between batch images (0: shared)
};


// d logL / d model of one pixel: logL = -1/2 mask (m - d)^2 / V(m) with
// V = max(m, 0)/t + bkg^2 (dV/dm = 1/t for m >= 0, the branch the reference's
// jnp.where(data >= 0, ...) takes) or the noise map's square (dV/dm = 0).
__device__ __forceinline__ double dlogl_dm(const ConvArgs &c, size_t p, double m, int img) {
  const size_t q = p + size_t(img) * size_t(c.dstride);  // per-image observation
  const double d = AT(c.data, q);
  double var, dvar;
  if (c.noise) {
    var = AT(c.noise, q) * AT(c.noise, q);
    dvar = 0.0;
  } else {
    const double it = c.expo ? rcp_refined(AT(c.expo, p)) : c.inv_t;  // 1 / exposure
    var = fma(m >= 0.0 ? m : 0.0, it, c.bkg2);
    dvar = m >= 0.0 ? it : 0.0;
  }
  const double r = (m - d) * rcp_refined(var);
  double gp = -(r - 0.5 * r * r * dvar);
  if (c.mask) gp *= AT(c.mask, p);
  return gp;
}

template <typename T>
__device__ __forceinline__ double chi2_pixel(const ConvArgs &c, size_t p, T model_t, int img) {
  const size_t q = p + size_t(img) * size_t(c.dstride);  // per-image observation
  const double m = double(model_t);
  const double d = AT(c.data, q);
  double var;
  if (c.noise) {
    const double nm = AT(c.noise, q);
    var = nm * nm;
  } else {
    const double it = c.expo ? rcp_refined(AT(c.expo, p)) : c.inv_t;  // 1 / exposure
    const double mpos = (m >= 0.0) ? m : 0.0;
    var = fma(mpos, it, c.bkg2);
  }
  double chi = (m - d) * (m - d) * rcp_refined(var);
  if (c.mask) chi *= AT(c.mask, p);
  return chi;
}

// out[y][x] = sum_{a,c} in[y - cy + a][x - cx + c] * Kf[a][c]  (Kf = flipped kernel).
// Block = TX columns x (TY / kWin) row groups; each thread computes kWin
// consecutive rows of one column from a register window that slides down the
// LDS tile, so consecutive lanes read consecutive LDS words.
// LW: the weights are staged in LDS behind the tile.  For launches of fewer
// blocks than CUs: there each step's scalar loads of the weights would miss
// the caches in turn (the kernel was just written by another launch), one
// memory latency per step.
template <typename T, int TY, bool LW>
__global__ __launch_bounds__(256) void conv_kernel(ConvArgs c, const T *__restrict__ in,
                                                   const double *__restrict__ kt, T *__restrict__ out,
                                                   double *__restrict__ partial) {
  constexpr int TX = ConvTile<T>::TX, NT = TX * TY / kWin;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tile = reinterpret_cast<T *>(smem_raw);
  const int img = blockIdx.y;
  const int tile_id = blockIdx.x;
  const int tyi = tile_id / c.n_tiles_x, txi = tile_id - tyi * c.n_tiles_x;
  const int y0 = tyi * TY, x0 = txi * TX;
  const int cy = (c.ky - 1) / 2, cx = (c.kx - 1) / 2;
  const int kyp = c.kyp;
  const int TH = TY + kyp - 1, TW = TX + c.kx - 1;
  const int THp = TH | 1;  // column stride of the column-major tile: odd, so lanes hit distinct banks
  const T *src = in + size_t(img) * c.nx * c.ny;
  const int tid = threadIdx.x;
  // kLd elements per thread and pass: all their loads are issued before the
  // first LDS store waits, so a launch of few blocks (LW) pays one memory
  // latency per pass rather than one per element.  Large launches hide that
  // latency across blocks and keep one element per pass (eight measured
  // slower for the float adjoint convolution of the gradient).
  constexpr int kLd = LW ? 8 : 1;
  for (int base = tid; base < TH * TW; base += NT * kLd) {
    T v[kLd];
    int at[kLd];
#pragma unroll
    for (int u = 0; u < kLd; ++u) {
      const int idx = base + u * NT;
      const int ly = idx / TW, lx = idx - ly * TW;
      int gy = y0 - cy + ly, gx = x0 - cx + lx;
      v[u] = T(0);
      at[u] = idx < TH * TW ? lx * THp + ly : -1;
      if (idx >= TH * TW || ly >= TY + c.ky - 1) {
        // past the tile, or rows that only meet the zero padding of the kernel height
      } else if (c.edge) {
        gy = gy < 0 ? 0 : (gy >= c.ny ? c.ny - 1 : gy);
        gx = gx < 0 ? 0 : (gx >= c.nx ? c.nx - 1 : gx);
        v[u] = AT(src, gy * c.nx + gx);
      } else if (gy >= 0 && gy < c.ny && gx >= 0 && gx < c.nx) {
        v[u] = AT(src, gy * c.nx + gx);
      }
    }
#pragma unroll
    for (int u = 0; u < kLd; ++u)
      if (at[u] >= 0) tile[at[u]] = v[u];
  }
  const int nch = kyp / kWin, nsteps = c.kx * nch;
  double *wl = reinterpret_cast<double *>(smem_raw + ((size_t(THp) * TW * sizeof(T) + 7) & ~size_t(7)));
  if constexpr (LW) {
    for (int idx = tid; idx < nsteps * kWin; idx += NT) wl[idx] = AT(kt, idx);
  } else {
    CHECK_SPAN(kt, size_t(nsteps + 1) * kWin * sizeof(double));
  }
  __syncthreads();
  const int lx = tid % TX, tg = tid / TX;
  // float64 accumulation in both forms (full-rate FMA on CDNA4)
  double acc[kWin];
#pragma unroll
  for (int r = 0; r < kWin; ++r) acc[r] = 0.0;
  // steps t = (kernel column cc, row chunk ch) flattened; the window of step
  // t + 1 is loaded while step t is computed
  const T *colbase = tile + lx * THp + tg * kWin;
  double buf[2 * kWin - 1];
#pragma unroll
  for (int q = 0; q < 2 * kWin - 1; ++q) buf[q] = double(colbase[q]);
  double w[kWin];
#pragma unroll
  for (int jj = 0; jj < kWin; ++jj) w[jj] = LW ? wl[jj] : kt[jj];
  int cc = 0, ch = 0;
  for (int t = 0; t < nsteps; ++t) {
    if (++ch == nch) {
      ch = 0;
      ++cc;
    }
    // next step's weights and window are requested before this step's FMAs
    // (the compiler still waits for them before the FMAs: the weights are
    // scalar loads, which complete out of order, so any wait on the shared
    // counter is a wait for all of them; latency is hidden across waves)
    double wn[kWin];
    T nb[2 * kWin - 1];
    if (t + 1 < nsteps) {
      if constexpr (LW) {
#pragma unroll
        for (int jj = 0; jj < kWin; ++jj) wn[jj] = wl[(t + 1) * kWin + jj];
      } else {
        const double *kw = kt + (t + 1) * kWin;  // = kt + cc * kyp + ch * kWin
#pragma unroll
        for (int jj = 0; jj < kWin; ++jj) wn[jj] = kw[jj];
      }
      const T *col = colbase + cc * THp + ch * kWin;
#pragma unroll
      for (int q = 0; q < 2 * kWin - 1; ++q) nb[q] = col[q];
    }
#pragma unroll
    for (int jj = 0; jj < kWin; ++jj) {
#pragma unroll
      for (int r = 0; r < kWin; ++r) acc[r] = fma(buf[r + jj], w[jj], acc[r]);
    }
#pragma unroll
    for (int q = 0; q < 2 * kWin - 1; ++q) buf[q] = double(nb[q]);
#pragma unroll
    for (int jj = 0; jj < kWin; ++jj) w[jj] = wn[jj];
  }
  const int x = x0 + lx;
  double chi = 0.0;
  if (x < c.nx) {
#pragma unroll
    for (int r = 0; r < kWin; ++r) {
      const int y = y0 + tg * kWin + r;
      if (y < c.ny) {
        const size_t p = size_t(y) * c.nx + x;
        const T m = T(acc[r]);
        if (out) AT(out, size_t(img) * c.nx * c.ny + p) = m;
        if (partial) chi += chi2_pixel<T>(c, p, m, img);
        if (c.gout) AT(static_cast<T *>(c.gout), size_t(img) * c.nx * c.ny + p) = T(dlogl_dm(c, p, double(m), img));
      }
    }
  }
  if (partial) {
    // wave reduction then one value per wave through LDS (fixed order)
    for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off, 64);
    __syncthreads();
    double *red = reinterpret_cast<double *>(smem_raw);
    const int wave = tid >> 6, lane = tid & 63;
    if (lane == 0) red[wave] = chi;
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      for (int w = 0; w < NT / 64; ++w) s += red[w];
      AT(partial, size_t(img) * gridDim.x + tile_id) = s;
    }
  }
}


// ------------------------------------------------ MFMA convolution (mixed) --
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// B fragments of the banded kernel rows, in the lane layout of
// v_mfma_f32_16x16x32_f16 (lane l holds B[8 (l / 16) + j][l % 16], j < 8),
// split into fp16 hi + lo of the kernel scaled by 2^sk (max |K| 2^sk in
// [2^14, 2^15)).  One thread per fragment; every block reduces the kernel
// maximum itself (K is small), block 0 stores sk in hdr[0].  The fragments of
// the padding steps (ky * nks up to n_steps) are zero.
// One 256-thread block (index b) of the fragment preparation.
__device__ void psf_mfma_prep_block(const double *__restrict__ k, int kx, int ky, int nks, int n_steps,
                                    int *__restrict__ hdr, half8 *__restrict__ frag, int b) {
  __shared__ double red[4];
  double m = 0.0;
  for (int i = threadIdx.x; i < kx * ky; i += blockDim.x) m = fmax(m, fabs(AT(k, i)));
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  const double mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const int sk = (mx > 0.0 && isfinite(mx)) ? 14 - ilogb(mx) : 0;
  if (b == 0 && threadIdx.x == 0) AT(hdr, 0) = sk;
  const double scale = ldexp(1.0, sk);  // exact power of two
  const int idx = b * 256 + threadIdx.x;
  if (idx >= n_steps * 2 * 64) return;
  const int lane = idx & 63, s = (idx >> 6) & 1, ak = idx >> 7;
  if (ak >= ky * nks) {
    AT(frag, idx) = half8{};
    return;
  }
  const int a = ak / nks, ks = ak - a * nks;
  const int n = lane & 15, g = lane >> 4;
  const double *krow = k + size_t(ky - 1 - a) * kx;
  half8 v;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int c = 32 * ks + 8 * g + j - n;  // tap of the flipped kernel row a
    const double w = (c >= 0 && c < kx) ? AT(krow, kx - 1 - c) * scale : 0.0;
    const _Float16 h = (_Float16)w;
    v[j] = s == 0 ? h : (_Float16)(w - (double)h);
  }
  AT(frag, idx) = v;
}

__global__ __launch_bounds__(256) void psf_mfma_prep_kernel(const double *__restrict__ k, int kx, int ky, int nks,
                                                            int n_steps, int *__restrict__ hdr,
                                                            half8 *__restrict__ frag) {
  psf_mfma_prep_block(k, kx, ky, nks, n_steps, hdr, frag, blockIdx.x);
}

// The PSF preparation of a call, folded into the record preparation's launch
// (the two are independent): mode 1 the flipped transposed kernel, 2 the MFMA
// fragments.
struct PsfJob {
  int mode;
  const double *k;
  int kx, ky, kyp, nks, n_steps;
  double *kt;
  int *hdr;
  half8 *frag;
};

int psf_job_blocks(const PsfJob &pj) {
  if (pj.mode == 1) return (pj.kyp * pj.kx + 255) / 256;
  if (pj.mode == 2) return (pj.n_steps * 2 * 64 + 255) / 256;
  if (pj.mode == 3) return (pj.n_steps * 64 + 255) / 256;
  return 0;
}

// B fragments of the float64 matrix-core convolution, [a][kk][lane] in the
// lane layout of v_mfma_f64_16x16x4_f64 (lane l holds B[l / 16][l % 16] of
// K-step kk): B[k][n] = Kf[a][k - n], Kf = the flipped kernel, zero outside;
// the n_steps - ky nk4 padding steps are zero.
__device__ void psf_mf64_prep_one(const double *__restrict__ k, int kx, int ky, int nk4, int n_steps,
                                  double *__restrict__ frag, int idx) {
  if (idx >= n_steps * 64) return;
  const int lane = idx & 63, step = idx >> 6;
  const int a = step / nk4, kk = step - a * nk4;
  const int c = 4 * kk + (lane >> 4) - (lane & 15);  // tap of the flipped row a
  AT(frag, idx) = (a < ky && c >= 0 && c < kx) ? AT(k, size_t(ky - 1 - a) * kx + (kx - 1 - c)) : 0.0;
}

// Blocks [0, nb_prep): one thread per (image, profile) record; the rest: the PSF job.
__global__ __launch_bounds__(256) void prep_psf_kernel(ModelArgs m, const double *__restrict__ params, int n_img,
                                                       double *__restrict__ recd, float *__restrict__ recf,
                                                       double *__restrict__ coefd, float *__restrict__ coeff,
                                                       double *__restrict__ dcoefd, float *__restrict__ dcoeff,
                                                       PsfJob pj, int nb_prep) {
  if (int(blockIdx.x) < nb_prep) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx < n_img * m.n_prof) prep_one(m, params, n_img, idx, recd, recf, coefd, coeff, dcoefd, dcoeff);
    return;
  }
  const int b = blockIdx.x - nb_prep;
  if (pj.mode == 2) {
    psf_mfma_prep_block(pj.k, pj.kx, pj.ky, pj.nks, pj.n_steps, pj.hdr, pj.frag, b);
  } else if (pj.mode == 3) {
    psf_mf64_prep_one(pj.k, pj.kx, pj.ky, pj.nks, pj.n_steps, pj.kt, b * 256 + threadIdx.x);
  } else {
    const int idx = b * 256 + threadIdx.x;
    if (idx < pj.kyp * pj.kx) psf_prep_one(pj.k, pj.kx, pj.ky, pj.kyp, pj.kt, 1, idx);
  }
}

// Direct 'same' convolution on the matrix cores: block = 64 rows x 16 TL
// columns (4 waves x TL tiles of 16 x 16; NKS K-steps of 32 per kernel row).
// The padded tile sits in LDS as fp16 hi and lo planes of the image scaled by
// 2^sa (sa from the per-image max); per kernel row a and K-step, three MFMAs
// (hi.hi, hi.lo, lo.hi) accumulate in float32, and the float sums are folded
// into float64 accumulators once per unrolled body of rows.
template <int NKS, int TL>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(1, 3))) void conv_mfma_kernel(
    ConvArgs c, int ls, const float *__restrict__ in,
                                                        const int *__restrict__ hdr,
                                                        const half8 *__restrict__ frag,
                                                        const float *__restrict__ blkmax, int nblk,
                                                        float *__restrict__ out, double *__restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // XCD-aware order: consecutive workgroups go to different XCDs (round
  // robin over 8), so the workgroups one XCD receives are given consecutive
  // (image, tile) indices and share their halos in that XCD's L2
  int img = blockIdx.y, tile_id = blockIdx.x;
  {
    const int n_blk = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
    if ((n_blk & 7) == 0) {
      const int logical = (lin & 7) * (n_blk >> 3) + (lin >> 3);
      img = logical / gridDim.x;
      tile_id = logical - img * gridDim.x;
    }
  }
  const int tyi = tile_id / c.n_tiles_x, txi = tile_id - tyi * c.n_tiles_x;
  const int Y0 = tyi * kMfTY, X0 = txi * 16 * TL;
  const int cy = (c.ky - 1) / 2, cx = (c.kx - 1) / 2;
  const int TH = kMfTY + c.ky - 1, TW = 16 * (TL - 1) + 32 * NKS;
  _Float16 *lh = reinterpret_cast<_Float16 *>(smem_raw);
  _Float16 *ll = lh + TH * ls;
  const int tid = threadIdx.x;
  const float *src = in + size_t(img) * c.nx * c.ny;
  // the wave index is the same on every lane: read it as a scalar, so the
  // tile rows a wave loads are addressed with scalar, not vector, arithmetic
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, li = lane & 15, lg = lane >> 4;
  // The padded tile: each wave fills rows wave, wave + 4, ..., lanes along a
  // row, in batches of kLdRows rows x kLdCols 64-column chunks.  Every load
  // reads a clamped (in-image) address and is zeroed after the fact outside
  // the image, so a whole batch is issued before the first load is waited
  // for; the first batch is in flight while the image maximum is reduced.
  constexpr int kLdRows = 8, kLdCols = (16 * (TL - 1) + 32 * NKS + 63) / 64;
  float v[kLdRows][kLdCols];
  auto load_batch = [&](int ly0) {
#pragma unroll
    for (int r = 0; r < kLdRows; ++r) {
      const int gy0 = Y0 - cy + ly0 + kMfWaves * r;
      const int gy = gy0 < 0 ? 0 : (gy0 >= c.ny ? c.ny - 1 : gy0);
      const bool row_ok = c.edge || gy0 == gy;
#pragma unroll
      for (int q = 0; q < kLdCols; ++q) {
        const int gx0 = X0 - cx + 64 * q + lane;
        const int gx = gx0 < 0 ? 0 : (gx0 >= c.nx ? c.nx - 1 : gx0);
        const float x = AT(src, size_t(gy) * c.nx + gx);
        v[r][q] = (row_ok && (c.edge || gx0 == gx)) ? x : 0.f;
      }
    }
  };
  load_batch(wave);
  // image max |value| from the render blocks' maxima
  __shared__ float smax[4];
  {
    float mv = 0.f;
    for (int b = tid; b < nblk; b += 256) mv = fmaxf(mv, AT(blkmax, size_t(img) * nblk + b));
    for (int off = 32; off > 0; off >>= 1) mv = fmaxf(mv, __shfl_xor(mv, off, 64));
    if ((tid & 63) == 0) smax[tid >> 6] = mv;
    __syncthreads();
  }
  const float mx = fmaxf(fmaxf(smax[0], smax[1]), fmaxf(smax[2], smax[3]));
  const int sa = (mx > 0.f && isfinite(mx)) ? 14 - ilogbf(mx) : 0;
  const int sk = AT(hdr, 0);
  const float scale_a = ldexpf(1.0f, sa);  // exact power of two
  auto store_batch = [&](int ly0) {
#pragma unroll
    for (int r = 0; r < kLdRows; ++r) {
      const int ly = ly0 + kMfWaves * r;
#pragma unroll
      for (int q = 0; q < kLdCols; ++q) {
        const int lx = 64 * q + lane;
        if (lx < TW && ly < TH) {
          const float xs = v[r][q] * scale_a;
          const _Float16 h = (_Float16)xs;
          lh[ly * ls + lx] = h;
          ll[ly * ls + lx] = (_Float16)(xs - (float)h);
        }
      }
    }
  };
  store_batch(wave);
  for (int ly0 = wave + kMfWaves * kLdRows; ly0 < TH; ly0 += kMfWaves * kLdRows) {
    load_batch(ly0);
    store_batch(ly0);
  }
  __syncthreads();
  double dacc[TL][4];
#pragma unroll
  for (int t = 0; t < TL; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) dacc[t][j] = 0.0;
  const int rowbase = (16 * wave + li) * ls + 8 * lg;
  // One step per (kernel row a, K block ks).  The rows are padded with
  // zero-B rows to a multiple of U; a body covers U rows (R = U * NKS steps).
  // Two register sets of B fragments (from L2) alternate between bodies: a
  // set is refilled with the steps 2R ahead as soon as its body's MFMAs are
  // issued, so its loads fly during the other set's body.  No load sits under
  // a branch (the refills past the last step re-read the last fragments and
  // are not used).  The float sums are folded into the float64 accumulators
  // once per pair of bodies (2U kernel rows).
  constexpr int U = mfma_unroll(NKS), R = U * NKS;
  const int ky_pad = (c.ky + U - 1) / U * U, n_steps = ky_pad * NKS;
  half8 bq[2][R][2];
  // The fragments are read as a buffer (resource in scalar registers): the
  // step's offset is a scalar and the lane's a fixed vector offset, so a load
  // needs no vector address arithmetic.  Reads past the n_steps fragments
  // (none are issued) would return zeros.
  const __amdgpu_buffer_rsrc_t frs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<half8 *>(frag), (short)0, n_steps * 2048, 0x00020000);
  const int voff = lane * 16;
  CHECK_SPAN(frag, size_t(n_steps) * 2048);
  auto refill = [&](half8 (&b)[R][2], int s0) {
#pragma unroll
    for (int p = 0; p < R; ++p) {
      const int q = __builtin_amdgcn_readfirstlane(s0 + p < n_steps ? s0 + p : n_steps - 1);
      b[p][0] = __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(frs, voff, q * 2048, 0));
      b[p][1] = __builtin_bit_cast(half8, __builtin_amdgcn_raw_buffer_load_b128(frs, voff, q * 2048 + 1024, 0));
    }
  };
  floatx4 acc[TL];
  auto zero = [&]() {
#pragma unroll
    for (int t = 0; t < TL; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  };
  auto fold = [&]() {
#pragma unroll
    for (int t = 0; t < TL; ++t)
#pragma unroll
      for (int j = 0; j < 4; ++j) dacc[t][j] += double(acc[t][j]);
  };
  auto body = [&](const half8 (&b)[R][2], int a0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int ar = a0 + u < c.ky ? a0 + u : c.ky - 1;  // padding rows re-read the last row (times zero B)
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        const int p = u * NKS + ks;
        half8 ah[TL], al[TL];
#pragma unroll
        for (int t = 0; t < TL; ++t) {
          const int off = rowbase + ar * ls + 16 * t + 32 * ks;
          ah[t] = *reinterpret_cast<const half8 *>(lh + off);
          al[t] = *reinterpret_cast<const half8 *>(ll + off);
        }
#pragma unroll
        for (int t = 0; t < TL; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[t], b[p][0], acc[t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < TL; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[t], b[p][1], acc[t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < TL; ++t)
          acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[t], b[p][0], acc[t], 0, 0, 0);
      }
    }
  };
  refill(bq[0], 0);
  refill(bq[1], R);
  int a0 = 0;
  for (; a0 + 2 * U <= ky_pad; a0 += 2 * U) {
    zero();
    body(bq[0], a0);
    refill(bq[0], (a0 + 2 * U) * NKS);
    body(bq[1], a0 + U);
    refill(bq[1], (a0 + 3 * U) * NKS);
    fold();
  }
  if (a0 < ky_pad) {  // an odd number of bodies: the last one
    zero();
    body(bq[0], a0);
    fold();
  }
  // lane l holds D[4 (l / 16) + j][l % 16] of each tile
  const double unscale = ldexp(1.0, -(sa + sk));
  // chi^2: one partial per 16 x 16 output tile, at [img][tile row][tile
  // column] of the image's 16 x 16 tiling, so the order of the sum does not
  // depend on the block width (which the plan picks by batch size): a request
  // has the same log-likelihood in any batch
  const int n16x = (c.nx + 15) >> 4, n16 = n16x * ((c.ny + 15) >> 4);
  const int ty16 = (Y0 >> 4) + wave;
#pragma unroll
  for (int t = 0; t < TL; ++t) {
    const int x = X0 + 16 * t + li;
    double chi = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int y = Y0 + 16 * wave + 4 * lg + j;
      if (x < c.nx && y < c.ny) {
        const size_t p = size_t(y) * c.nx + x;
        const float m = float(dacc[t][j] * unscale);
        if (out) AT(out, size_t(img) * c.nx * c.ny + p) = m;
        if (partial) chi += chi2_pixel<float>(c, p, m, img);
        if (c.gout)
          AT(static_cast<float *>(c.gout), size_t(img) * c.nx * c.ny + p) = float(dlogl_dm(c, p, double(m), img));
      }
    }
    if (partial) {
      for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off, 64);
      const int tx16 = (X0 >> 4) + t;
      // tiles past the image's edge hold no pixel and have no slot
      if (lane == 0 && 16 * tx16 < c.nx && 16 * ty16 < c.ny)
        AT(partial, size_t(img) * n16 + size_t(ty16) * n16x + tx16) = chi;
    }
  }
}


// Direct 'same' convolution in float64 on the matrix cores (FP64 form):
// block = 64 rows x 16 TL columns (4 waves x TL tiles of 16 x 16), per
// kernel row a and K-step kk one v_mfma_f64_16x16x4_f64 per tile; the padded
// tile sits in LDS, the B fragments (psf_mf64_prep_one) come from L2 through
// buffer loads.  Same outputs, chi^2 partials (per 16 x 16 tile) and
// dlogL/dmodel as conv_mfma_kernel, all in float64.
typedef double double4_t __attribute__((ext_vector_type(4)));

template <int TL>
__global__ __launch_bounds__(256) void conv_mf64_kernel(ConvArgs c, int ls, int nk4, const double *__restrict__ in,
                                                        const double *__restrict__ frag, double *__restrict__ out,
                                                        double *__restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double *tile = reinterpret_cast<double *>(smem_raw);
  // XCD-aware order, as conv_mfma_kernel
  int img = blockIdx.y, tile_id = blockIdx.x;
  {
    const int n_blk = gridDim.x * gridDim.y, lin = blockIdx.y * gridDim.x + blockIdx.x;
    if ((n_blk & 7) == 0) {
      const int logical = (lin & 7) * (n_blk >> 3) + (lin >> 3);
      img = logical / gridDim.x;
      tile_id = logical - img * gridDim.x;
    }
  }
  const int tyi = tile_id / c.n_tiles_x, txi = tile_id - tyi * c.n_tiles_x;
  const int Y0 = tyi * kMfTY, X0 = txi * 16 * TL;
  const int cy = (c.ky - 1) / 2, cx = (c.kx - 1) / 2;
  const int TH = kMfTY + c.ky - 1, TW = 16 * (TL - 1) + 4 * nk4;
  const int tid = threadIdx.x;
  const double *src = in + size_t(img) * c.nx * c.ny;
  // padded tile: eight loads in flight per thread before the LDS stores
  constexpr int kLd = 8;
  for (int base = tid; base < TH * TW; base += 256 * kLd) {
    double v[kLd];
    int at[kLd];
#pragma unroll
    for (int u = 0; u < kLd; ++u) {
      const int idx = base + u * 256;
      const int ly = idx / TW, lx = idx - ly * TW;
      int gy = Y0 - cy + ly, gx = X0 - cx + lx;
      v[u] = 0.0;
      at[u] = idx < TH * TW ? ly * ls + lx : -1;
      if (idx >= TH * TW) {
      } else if (c.edge) {
        gy = gy < 0 ? 0 : (gy >= c.ny ? c.ny - 1 : gy);
        gx = gx < 0 ? 0 : (gx >= c.nx ? c.nx - 1 : gx);
        v[u] = AT(src, size_t(gy) * c.nx + gx);
      } else if (gy >= 0 && gy < c.ny && gx >= 0 && gx < c.nx) {
        v[u] = AT(src, size_t(gy) * c.nx + gx);
      }
    }
#pragma unroll
    for (int u = 0; u < kLd; ++u)
      if (at[u] >= 0) tile[at[u]] = v[u];
  }
  __syncthreads();
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63, li = lane & 15, lg = lane >> 4;
  // lane (li, lg) reads A[li][lg] of each K-step: tile row 16 wave + li + a,
  // column 16 t + 4 kk + lg
  const double *arow = tile + (16 * wave + li) * ls + lg;
  // n_steps: a multiple of kMf64B; the steps past ky nk4 have zero B
  // fragments and re-read A of the last kernel row (a stays at ky - 1)
  const int n_steps = (c.ky * nk4 + kMf64B - 1) / kMf64B * kMf64B;
  const __amdgpu_buffer_rsrc_t frs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(frag), (short)0, n_steps * 512, 0x00020000);
  const int voff = lane * 8;
  CHECK_SPAN(frag, size_t(n_steps) * 512);
  double4_t acc[TL];
#pragma unroll
  for (int t = 0; t < TL; ++t) acc[t] = double4_t{0.0, 0.0, 0.0, 0.0};
  // the B fragments of kMf64B steps are requested ahead of their MFMAs
  double bq[kMf64B];
#pragma unroll
  for (int p = 0; p < kMf64B; ++p)
    bq[p] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(frs, voff, p * 512, 0));
  // A of a step is read from LDS one step ahead, so its latency falls under
  // the previous step's MFMAs (a block's LDS leaves room for one block per CU
  // at the usual tile widths: no other wave hides it)
  int a = 0, kk = 0;
  double av[TL];
#pragma unroll
  for (int t = 0; t < TL; ++t) av[t] = arow[16 * t];
  for (int s0 = 0; s0 < n_steps; s0 += kMf64B) {
#pragma unroll
    for (int p = 0; p < kMf64B; ++p) {
      int a1 = a, kk1 = kk + 1;
      if (kk1 == nk4) {
        kk1 = 0;
        if (a1 + 1 < c.ky) ++a1;
      }
      const double *apn = arow + a1 * ls + 4 * kk1;
      double an[TL];
#pragma unroll
      for (int t = 0; t < TL; ++t) an[t] = apn[16 * t];
#pragma unroll
      for (int t = 0; t < TL; ++t) acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[t], bq[p], acc[t], 0, 0, 0);
#pragma unroll
      for (int t = 0; t < TL; ++t) av[t] = an[t];
      a = a1;
      kk = kk1;
      // refill slot p with step s0 + p + kMf64B (the refills past the end re-read the last step, unused)
      const int q = __builtin_amdgcn_readfirstlane(s0 + p + kMf64B < n_steps ? s0 + p + kMf64B : n_steps - 1);
      bq[p] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(frs, voff, q * 512, 0));
    }
  }
  // result register j of lane l holds D[l / 16 + 4 j][l % 16] of each tile
  // (measured on gfx950); chi^2 partials per 16 x 16 tile
  const int n16x = (c.nx + 15) >> 4, n16 = n16x * ((c.ny + 15) >> 4);
  const int ty16 = (Y0 >> 4) + wave;
#pragma unroll
  for (int t = 0; t < TL; ++t) {
    const int x = X0 + 16 * t + li;
    double chi = 0.0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int y = Y0 + 16 * wave + lg + 4 * j;
      if (x < c.nx && y < c.ny) {
        const size_t p = size_t(y) * c.nx + x;
        const double m = acc[t][j];
        if (out) AT(out, size_t(img) * c.nx * c.ny + p) = m;
        if (partial) chi += chi2_pixel<double>(c, p, m, img);
        if (c.gout) AT(static_cast<double *>(c.gout), size_t(img) * c.nx * c.ny + p) = dlogl_dm(c, p, m, img);
      }
    }
    if (partial) {
      for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off, 64);
      const int tx16 = (X0 >> 4) + t;
      if (lane == 0 && 16 * tx16 < c.nx && 16 * ty16 < c.ny)
        AT(partial, size_t(img) * n16 + size_t(ty16) * n16x + tx16) = chi;
    }
  }
}

// ------------------------------------------------------------- gradient ---
// Number of kwargs of a profile type (its slots in a parameter block).
__host__ __device__ constexpr int n_params_of(int type) {
  switch (type) {
    case LENSIM_EPL: return 6;
    case LENSIM_SIE: return 5;
    case LENSIM_NIE: return 6;
    case LENSIM_NFW: return 4;
    case LENSIM_TNFW: return 5;
    case LENSIM_SHEAR: return 4;
    case LENSIM_CONVERGENCE: return 3;
    case LENSIM_SERSIC: return 5;
    case LENSIM_SERSIC_ELLIPSE: return 7;
    case LENSIM_GAUSSIAN: return 4;
    default: return 0;
  }
}

struct GradArgs {
  const double *params;   // [n_img][n_prof][8]
  const double *coefd;    // EPL F_n (double)
  const float *coeff;     // EPL F_n (float)
  const double *dcoefd;   // EPL dF_n/dt
  const float *dcoeff;
  const void *glow;       // [n_img][ny][nx] dlogL / d(unconvolved image), compute type
  double *gpart;          // [n_img][n_prof][nblk][8]
  void *gbeta;            // [n_img][ny][nx][s][s][2] dS/dbeta of the sources' sum, compute type, or NULL
};

// A light profile's flux and its derivatives in its own kwargs (the N seeds
// of ps) at (x, y): closed form for SERSIC_ELLIPSE, duals otherwise.
template <typename T, int N>
__device__ __forceinline__ gm::Dual<T, N> light_grad(int type, const gm::Dual<T, N> *ps, const double *pk, T x,
                                                     T y) {
  using S = gm::Dual<T, N>;
  if constexpr (N == 7) {
    if (type == LENSIM_SERSIC_ELLIPSE) {
      T gp[7], gx, gy;
      S r;
      r.v = gm::sersic_ellipse_grad<T>(pk, x, y, gp, gx, gy);
#pragma unroll
      for (int i = 0; i < 7; ++i) r.d[i] = gp[i];
      return r;
    }
  }
  return gm::light_g<T, S>(type, ps, S(x), S(y));
}

// Sum of the source light profiles at beta and its gradient in beta (the
// tangents d[0], d[1]): closed form for SERSIC_ELLIPSE, two dual seeds
// otherwise.
template <typename T>
__device__ __forceinline__ gm::Dual<T, 2> sources_at(const ModelArgs &m, const GradArgs &ga, int img, T bx, T by) {
  using S2 = gm::Dual<T, 2>;
  const S2 bx2 = S2::seed(bx, 0), by2 = S2::seed(by, 1);
  S2 f2 = S2(T(0));
  for (int j = m.n_defl; j < m.n_defl + m.n_source; ++j) {
    const double *pj = ga.params + (size_t(img) * m.n_prof + j) * LENSIM_NPARAM;
    CHECK_SPAN(pj, LENSIM_NPARAM * sizeof(double));
    const int tj = m.types[j - m.n_bulk];
    if (tj == LENSIM_SERSIC_ELLIPSE) {
      T gp[7], gx, gy;
      f2.v += gm::sersic_ellipse_grad<T>(pj, bx, by, gp, gx, gy);
      f2.d[0] += gx;
      f2.d[1] += gy;
      continue;
    }
    S2 pl[LENSIM_NPARAM];
#pragma unroll
    for (int i = 0; i < LENSIM_NPARAM; ++i) pl[i] = S2(T(pj[i]));
    f2 = f2 + gm::light_g<T, S2>(tj, pl, bx2, by2);
  }
  return f2;
}

// Contribution of one ray to d flux / d(kwargs of profile k), with the
// profile's own kwargs as N dual seeds; w = dlogL / d(ray flux).
// gb: this ray's dS/dbeta of the sources' sum (two values), when the model
// has lenses and sources: the first source's launch writes it, the lens
// launches (which run after it) read it instead of deflecting by the other
// deflectors and evaluating the sources again.
template <typename T, int N, bool HALO>
__device__ __forceinline__ void ray_grad(const ModelArgs &m, const GradArgs &ga, const RecPtrs<T> &rp, int img,
                                         int k, int type, double x, double y, double w, double *acc, T *gb) {
  using S = gm::Dual<T, N>;
  const double *pk = ga.params + (size_t(img) * m.n_prof + k) * LENSIM_NPARAM;
  CHECK_SPAN(pk, LENSIM_NPARAM * sizeof(double));
  S ps[LENSIM_NPARAM];
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) ps[i] = i < N ? S::seed(T(pk[i]), i) : S(T(pk[i]));
  S flux;
  if (k < m.n_lens && gb) {
    // d flux = dS/dbeta . d beta, d beta = -alpha_scaling d alpha_k
    S ax = S(T(0)), ay = S(T(0));
    const int e = m.epl_index[k];
    const size_t co = (e >= 0) ? (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX : 0;
    int nser = 1;
    if (type == LENSIM_EPL)
      nser = int(sizeof(T) == sizeof(float) ? rp.recd[k * kRec + EPL_NF] : rp.recd[k * kRec + EPL_NMAX]);
    const T *c = sizeof(T) == sizeof(float) ? reinterpret_cast<const T *>(ga.coeff + co)
                                            : reinterpret_cast<const T *>(ga.coefd + co);
    const T *dc = sizeof(T) == sizeof(float) ? reinterpret_cast<const T *>(ga.dcoeff + co)
                                             : reinterpret_cast<const T *>(ga.dcoefd + co);
    if (e >= 0) {
      CHECK_SPAN(c, LENSIM_EPL_NMAX * sizeof(T));
      CHECK_SPAN(dc, LENSIM_EPL_NMAX * sizeof(T));
    }
    gm::defl_g<T, S>(type, ps, c, dc, nser, S(T(x)), S(T(y)), ax, ay);
    const T gx = AT(gb, 0), gy = AT(gb, 1), as = T(m.alpha_scaling);
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] += w * double(-as * (gx * ax.d[i] + gy * ay.d[i]));
    return;
  }
  if (k < m.n_lens) {
    // d/d(lens kwargs): the other deflectors in the forward form, this one in duals
    double axp = 0.0, ayp = 0.0;
    for (int j = 0; j < m.n_lens; ++j) {
      if (j == k) continue;
      const int e = m.epl_index[j];
      const size_t co = (e >= 0) ? (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX : 0;
      add_deflection<T, HALO>(m.types[j], Rec<T>{rp.recd + j * kRec, rp.rect + j * kRec}, rp.coefd + co,
                              rp.coeft + co, x, y, axp, ayp);
    }
    deflect_bulk<T, HALO>(m, rp, x, y, axp, ayp);
    S ax = S(T(axp)), ay = S(T(ayp));
    const int e = m.epl_index[k];
    const size_t co = (e >= 0) ? (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX : 0;
    int nser = 1;
    if (type == LENSIM_EPL)
      nser = int(sizeof(T) == sizeof(float) ? rp.recd[k * kRec + EPL_NF] : rp.recd[k * kRec + EPL_NMAX]);
    const T *c = sizeof(T) == sizeof(float) ? reinterpret_cast<const T *>(ga.coeff + co)
                                            : reinterpret_cast<const T *>(ga.coefd + co);
    const T *dc = sizeof(T) == sizeof(float) ? reinterpret_cast<const T *>(ga.dcoeff + co)
                                             : reinterpret_cast<const T *>(ga.dcoefd + co);
    if (e >= 0) {
      CHECK_SPAN(c, LENSIM_EPL_NMAX * sizeof(T));
      CHECK_SPAN(dc, LENSIM_EPL_NMAX * sizeof(T));
    }
    const S xs = S(T(x)), ys = S(T(y));
    gm::defl_g<T, S>(type, ps, c, dc, nser, xs, ys, ax, ay);
    const S bx = xs - ax * T(m.alpha_scaling), by = ys - ay * T(m.alpha_scaling);
    // the sources' gradient in beta, chained into this deflector's kwargs:
    // d flux = dS/dbeta_x d beta_x + dS/dbeta_y d beta_y
    const gm::Dual<T, 2> f2 = sources_at<T>(m, ga, img, bx.v, by.v);
    flux.v = f2.v;
#pragma unroll
    for (int i = 0; i < N; ++i) flux.d[i] = f2.d[0] * bx.d[i] + f2.d[1] * by.d[i];
  } else if (k < m.n_defl + m.n_source) {
    // d/d(source kwargs) at the deflected position
    T bxs, bys;
    if (m.mp_n) {  // multi-plane: mp_grad_kernel stored beta at every ray
      bxs = AT(gb, 0);
      bys = AT(gb, 1);
    } else {
      double ax = 0.0, ay = 0.0;
      deflect_all<T, HALO>(m, rp, img, x, y, ax, ay);
      bxs = T(x - ax * m.alpha_scaling);
      bys = T(y - ay * m.alpha_scaling);
    }
    bool done = false;
    if (gb && !m.mp_n) {  // the first source's launch: also dS/dbeta of the sources' sum for the lens launches
      if constexpr (N == 7) {
        if (m.n_source == 1 && type == LENSIM_SERSIC_ELLIPSE) {  // one evaluation gives both
          T gp[7], gx, gy;
          flux.v = gm::sersic_ellipse_grad<T>(pk, bxs, bys, gp, gx, gy);
#pragma unroll
          for (int i = 0; i < 7; ++i) flux.d[i] = gp[i];
          AT(gb, 0) = gx;
          AT(gb, 1) = gy;
          done = true;
        }
      }
      if (!done) {
        const gm::Dual<T, 2> f2 = sources_at<T>(m, ga, img, bxs, bys);
        AT(gb, 0) = f2.d[0];
        AT(gb, 1) = f2.d[1];
      }
    }
    if (!done) flux = light_grad<T, N>(type, ps, pk, bxs, bys);
  } else {
    // d/d(lens light kwargs) at the ray
    flux = light_grad<T, N>(type, ps, pk, T(x), T(y));
  }
#pragma unroll
  for (int i = 0; i < N; ++i) acc[i] += w * double(flux.d[i]);
}

// ray_grad for an inline NFW or TNFW deflector k: the other deflectors in the
// forward form, this one's duals in double in both forms (the float closed
// forms of the NFW family cancel, as for the bulk list's duals).
template <typename T, int N, bool HALO>
__device__ __forceinline__ void halo_ray_grad(const ModelArgs &m, const GradArgs &ga, const RecPtrs<T> &rp, int img,
                                              int k, int type, double x, double y, double w, double *acc,
                                              const T *gb) {
  using S = gm::Dual<double, N>;
  const double *pk = ga.params + (size_t(img) * m.n_prof + k) * LENSIM_NPARAM;
  CHECK_SPAN(pk, LENSIM_NPARAM * sizeof(double));
  if (gb) {  // dS/dbeta from the first source's launch
    S ps[LENSIM_NPARAM];
#pragma unroll
    for (int i = 0; i < LENSIM_NPARAM; ++i) ps[i] = i < N ? S::seed(pk[i], i) : S(pk[i]);
    S ax = S(0.0), ay = S(0.0);
    gm::defl_g<double, S>(type, ps, static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), 1,
                          S(x), S(y), ax, ay);
    const double gx = double(AT(gb, 0)), gy = double(AT(gb, 1)), as = m.alpha_scaling;
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] += w * (-as * (gx * ax.d[i] + gy * ay.d[i]));
    return;
  }
  double axp = 0.0, ayp = 0.0;
  for (int j = 0; j < m.n_lens; ++j) {
    if (j == k) continue;
    const int e = m.epl_index[j];
    const size_t co = (e >= 0) ? (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX : 0;
    add_deflection<T, HALO>(m.types[j], Rec<T>{rp.recd + j * kRec, rp.rect + j * kRec}, rp.coefd + co,
                            rp.coeft + co, x, y, axp, ayp);
  }
  deflect_bulk<T, HALO>(m, rp, x, y, axp, ayp);
  S ps[LENSIM_NPARAM];
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) ps[i] = i < N ? S::seed(pk[i], i) : S(pk[i]);
  S ax = S(axp), ay = S(ayp);
  gm::defl_g<double, S>(type, ps, static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), 1,
                        S(x), S(y), ax, ay);
  const S bx = S(x) - ax * m.alpha_scaling, by = S(y) - ay * m.alpha_scaling;
  const gm::Dual<T, 2> f2 = sources_at<T>(m, ga, img, T(bx.v), T(by.v));
#pragma unroll
  for (int i = 0; i < N; ++i) acc[i] += w * (double(f2.d[0]) * bx.d[i] + double(f2.d[1]) * by.d[i]);
}

// The deflection of the EPL deflector k at (x, y) and its derivatives by the
// six kwargs (theta_E, gamma, e1, e2, center_x, center_y).  Two dual seeds
// carry gamma and the axis ratio q; the rest is in closed form:
//  * alpha scales as b^t with b = theta_E sqrt(q), so d alpha / d theta_E =
//    t alpha / theta_E;
//  * d alpha / d center = -H, H = d alpha / dx the Hessian, written as the
//    reference's EPLMajorAxis.hessian (LensModel/Profiles/epl.py:398-440, and
//    the rotation of :203-232): kappa = (2 - t)/2 (b/R)^t and the shear of
//    its eq. (17) from alpha itself;
//  * alpha(x) = R(phi) alpha_M(R(phi)^T (x - c)), so d alpha / d phi =
//    J alpha - H J (x - c) with J the quarter turn;
//  * (e1, e2) -> (phi, q) by the closed forms of Util/param_util.py:88-100
//    (phi = atan2(e2, e1) / 2, c = min(|e|, 0.9999), q = (1 - c) / (1 + c);
//    dq = 0 where c is clamped).
template <typename T>
__device__ __forceinline__ void epl_dalpha(const ModelArgs &m, const GradArgs &ga, const RecPtrs<T> &rp, int img,
                                           int k, const double *pk, double x, double y, T &a_x, T &a_y, T *dax,
                                           T *day) {
  using S = gm::Dual<T, 2>;
  S ps[LENSIM_NPARAM];
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) ps[i] = i == 1 ? S::seed(T(pk[1]), 0) : S(T(pk[i]));
  const int e = m.epl_index[k];
  const size_t co = (size_t(img) * m.n_epl + e) * LENSIM_EPL_NMAX;
  const int nser = int(sizeof(T) == sizeof(float) ? rp.recd[k * kRec + EPL_NF] : rp.recd[k * kRec + EPL_NMAX]);
  const T *c = sizeof(T) == sizeof(float) ? reinterpret_cast<const T *>(ga.coeff + co)
                                          : reinterpret_cast<const T *>(ga.coefd + co);
  const T *dc = sizeof(T) == sizeof(float) ? reinterpret_cast<const T *>(ga.dcoeff + co)
                                           : reinterpret_cast<const T *>(ga.dcoefd + co);
  CHECK_SPAN(c, LENSIM_EPL_NMAX * sizeof(T));
  CHECK_SPAN(dc, LENSIM_EPL_NMAX * sizeof(T));
  // q, cos phi and sin phi from the record, not per ray
  const double *rdk = rp.recd + k * kRec;
  const T q = T(rdk[EPL_Q]), cs = T(rdk[EPL_C]), sn = T(rdk[EPL_S]);
  S ax = S(T(0)), ay = S(T(0));
  gm::epl_g_geo<T>(ps, S::seed(q, 1), S(cs), S(sn), c, dc, nser, S(T(x)), S(T(y)), ax, ay);
  a_x = ax.v;
  a_y = ay.v;
  const T thetaE = T(pk[0]), t = T(pk[1]) - T(1);
  const T dth = thetaE != T(0) ? t / thetaE : T(0);
  dax[0] = a_x * dth;
  day[0] = a_y * dth;
  dax[1] = ax.d[0];
  day[1] = ay.d[0];
  // Hessian in the major-axis frame, rotated back
  const T dx = T(x) - T(pk[4]), dy = T(y) - T(pk[5]);
  const T xr = dx * cs + dy * sn, yr = dy * cs - dx * sn;
  const T r = gm::sqrt(xr * xr + yr * yr);
  T fxx = T(0), fxy = T(0), fyy = T(0);
  if (r > T(0)) {
    T R = gm::sqrt(q * q * xr * xr + yr * yr);
    R = R >= T(0.00000001) ? R : T(0.00000001);
    const T kappa = (T(2) - t) / T(2) * gm::exp(t * (T(rdk[EPL_LNB]) - gm::log(R)));  // (b/R)^t
    const T ct = xr / r, st = yr / r;
    const T c2 = ct * ct * T(2) - T(1), s2 = st * ct * T(2);
    const T fr = a_x * cs + a_y * sn, fi = a_y * cs - a_x * sn;  // alpha in the major-axis frame
    const T g1 = (T(1) - t) * (fr * ct - fi * st) / r - kappa * c2;
    const T g2 = (T(1) - t) * (fi * ct + fr * st) / r - kappa * s2;
    const T c2p = cs * cs - sn * sn, s2p = T(2) * sn * cs;
    const T G1 = c2p * g1 - s2p * g2, G2 = s2p * g1 + c2p * g2;
    fxx = kappa + G1;
    fyy = kappa - G1;
    fxy = G2;
  }
  // orientation: d alpha / d phi = J alpha - H J (x - c)
  const T dphx = fxx * dy - fxy * dx - a_y, dphy = fxy * dy - fyy * dx + a_x;
  const T e1 = T(pk[2]), e2 = T(pk[3]);
  const T c2 = e1 * e1 + e2 * e2, ce = gm::sqrt(c2);
  const T h = T(0.5) / c2;  // d phi / d(e1, e2) = (-e2, e1) / (2 c^2)
  const T dp1 = -e2 * h, dp2 = e1 * h;
  const T dqdc = ce < T(0.9999) ? T(-2) / ((T(1) + ce) * (T(1) + ce)) : T(0);
  const T dq1 = dqdc * e1 / ce, dq2 = dqdc * e2 / ce;
  dax[2] = ax.d[1] * dq1 + dphx * dp1;
  day[2] = ay.d[1] * dq1 + dphy * dp1;
  dax[3] = ax.d[1] * dq2 + dphx * dp2;
  day[3] = ay.d[1] * dq2 + dphy * dp2;
  dax[4] = -fxx;
  day[4] = -fxy;
  dax[5] = -fxy;
  day[5] = -fyy;
}

// The deflection of the SIE / NIE deflector k at (x, y) and its derivatives by
// its kwargs (theta_E, e1, e2[, s_scale], center_x, center_y).  Dual seeds
// carry q, s_scale and the centre; the axis ratio and orientation come from
// the ellipticity (q, by Util/param_util.py:88-100) and the record (cos phi,
// sin phi), not per-ray dual atan2 / cos / sin; alpha is linear in theta_E
// (b = theta_E sqrt(q)); d alpha / d phi = J alpha - H J (x - c) as for the
// EPL, H = -(d alpha / d center).
template <typename T, int TYPE>
__device__ __forceinline__ void nie_dalpha(const double *rd, const double *pk, double x, double y, T &a_x, T &a_y,
                                           T *dax, T *day) {
  using S = gm::Dual<T, 4>;  // q, s_scale, center_x, center_y
  constexpr bool kNie = TYPE == LENSIM_NIE;
  const T e1 = T(pk[1]), e2 = T(pk[2]);
  const T c2 = e1 * e1 + e2 * e2, ce = gm::sqrt(c2);
  const T cc = ce < T(0.9999) ? ce : T(0.9999);
  const T q = (T(1) - cc) / (T(1) + cc);
  const T s_scale = kNie ? T(pk[3]) : T(0.0000000001);
  const T cx = T(pk[kNie ? 4 : 3]), cy = T(pk[kNie ? 5 : 4]);
  S ax = S(T(0)), ay = S(T(0));
  gm::nie_g_geo<T>(S(T(pk[0])), S::seed(q, 0), S(T(rd[NIE_C])), S(T(rd[NIE_S])), S::seed(s_scale, 1),
                   S::seed(cx, 2), S::seed(cy, 3), S(T(x)), S(T(y)), ax, ay);
  a_x = ax.v;
  a_y = ay.v;
  const T thetaE = T(pk[0]);
  const T ith = thetaE != T(0) ? T(1) / thetaE : T(0);
  dax[0] = a_x * ith;
  day[0] = a_y * ith;
  // H = d alpha / d(x, y) = -(d alpha / d center)
  const T hxx = -ax.d[2], hxy = -ax.d[3], hyx = -ay.d[2], hyy = -ay.d[3];
  const T dx = T(x) - cx, dy = T(y) - cy;
  const T dphx = hxx * dy - hxy * dx - a_y, dphy = hyx * dy - hyy * dx + a_x;
  const T h = T(0.5) / c2;  // d phi / d(e1, e2) = (-e2, e1) / (2 c^2)
  const T dp1 = -e2 * h, dp2 = e1 * h;
  const T dqdc = ce < T(0.9999) ? T(-2) / ((T(1) + ce) * (T(1) + ce)) : T(0);
  const T dq1 = dqdc * e1 / ce, dq2 = dqdc * e2 / ce;
  dax[1] = ax.d[0] * dq1 + dphx * dp1;
  day[1] = ay.d[0] * dq1 + dphy * dp1;
  dax[2] = ax.d[0] * dq2 + dphx * dp2;
  day[2] = ay.d[0] * dq2 + dphy * dp2;
  constexpr int kc = kNie ? 4 : 3;  // the centre's slots
  if constexpr (kNie) {
    dax[3] = ax.d[1];
    day[3] = ay.d[1];
  }
  dax[kc] = ax.d[2];
  day[kc] = ay.d[2];
  dax[kc + 1] = ax.d[3];
  day[kc + 1] = ay.d[3];
}

// Contribution of one ray to d flux / d(kwargs of the EPL, SIE or NIE deflector k).
template <typename T, bool HALO, int TYPE = LENSIM_EPL>
__device__ __forceinline__ void lens_ray_grad(const ModelArgs &m, const GradArgs &ga, const RecPtrs<T> &rp, int img,
                                             int k, double x, double y, double w, double *acc, const T *gb) {
  const double *pk = ga.params + (size_t(img) * m.n_prof + k) * LENSIM_NPARAM;
  CHECK_SPAN(pk, LENSIM_NPARAM * sizeof(double));
  double axp = 0.0, ayp = 0.0;  // the other deflectors in the forward form (not needed with gb)
  if (!gb) {
    for (int j = 0; j < m.n_lens; ++j) {
      if (j == k) continue;
      const int ej = m.epl_index[j];
      const size_t coj = (ej >= 0) ? (size_t(img) * m.n_epl + ej) * LENSIM_EPL_NMAX : 0;
      add_deflection<T, HALO>(m.types[j], Rec<T>{rp.recd + j * kRec, rp.rect + j * kRec}, rp.coefd + coj,
                              rp.coeft + coj, x, y, axp, ayp);
    }
    deflect_bulk<T, HALO>(m, rp, x, y, axp, ayp);
  }
  T a_x, a_y, dax[6], day[6];
  if constexpr (TYPE == LENSIM_EPL)
    epl_dalpha<T>(m, ga, rp, img, k, pk, x, y, a_x, a_y, dax, day);
  else
    nie_dalpha<T, TYPE>(rp.recd + k * kRec, pk, x, y, a_x, a_y, dax, day);
  constexpr int N = n_params_of(TYPE);
  const T as = T(m.alpha_scaling);
  T gx, gy;
  if (gb) {
    gx = AT(gb, 0);
    gy = AT(gb, 1);
  } else {
    const T bxv = T(x) - (T(axp) + a_x) * as, byv = T(y) - (T(ayp) + a_y) * as;
    const gm::Dual<T, 2> f2 = sources_at<T>(m, ga, img, bxv, byv);
    gx = f2.d[0];
    gy = f2.d[1];
  }
#pragma unroll
  for (int i = 0; i < N; ++i) acc[i] += w * double(-as * (gx * dax[i] + gy * day[i]));
}

// One block = kGradPix x 256 pixels x one image of profile k, whose type
// TYPE is fixed at compile time (one launch per profile: each type's kernel
// keeps only its own registers and code); each thread takes kGradPix pixels
// 256 apart, so the block reduction and the per-thread set-up are shared by
// kGradPix pixels' rays.
// The float form is held to 128 VGPRs (four waves per SIMD instead of three
// for the EPL's duals; no spills): 239k -> 250k gradients/s.  The FP64 form,
// and the NFW / TNFW kernels with their double duals, would spill.
template <typename T, int TYPE>
constexpr int grad_waves_per_eu() {
  return (sizeof(T) == sizeof(float) && TYPE != LENSIM_NFW && TYPE != LENSIM_TNFW) ? 4 : 1;
}

template <typename T, bool HALO, int TYPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(grad_waves_per_eu<T, TYPE>()))) void
grad_kernel(ModelArgs m, GridArgs g, GradArgs ga,
                                                   const double *__restrict__ recd, const float *__restrict__ recf,
                                                   const double *__restrict__ coefd, const float *__restrict__ coeff,
                                                   int k) {
  const int img = blockIdx.y;
  const int npix = g.nx * g.ny;
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  constexpr int type = TYPE;
  T *gbeta = static_cast<T *>(ga.gbeta);
  double acc[LENSIM_NPARAM];
#pragma unroll
  for (int q = 0; q < LENSIM_NPARAM; ++q) acc[q] = 0.0;
  for (int pp = 0; pp < kGradPix; ++pp) {
  const int pix0 = (blockIdx.x * kGradPix + pp) * blockDim.x + threadIdx.x;
  const bool active = pix0 < npix;
  const int pix = active ? pix0 : npix - 1;
  int i;
  const int j = div_by(pix, g.nx, g.inv_nx, i);
  // the pixel's flux is scale * flux_scaling * (sum over its rays)
  const double w = active ? double(AT(static_cast<const T *>(ga.glow), size_t(img) * npix + pix)) * g.scale *
                                m.flux_scaling
                          : 0.0;
  const double ra_c = i * g.m00 + j * g.m01 + g.ra0;
  const double dec_c = i * g.m10 + j * g.m11 + g.dec0;
  const double inv_s = g.inv_s;
  const double o0 = g.o0;
  // a lens profile only matters if a source is there to be lensed
  const bool live = active && w != 0.0 && (k >= m.n_defl || m.n_source > 0) && evaluated(g, i, j);
  for (int sy = 0; live && sy < g.s; ++sy) {
    const double oy = o0 + sy * inv_s;
    for (int sx = 0; sx < g.s; ++sx) {
      const double ox = o0 + sx * inv_s;
      const double x = ra_c + g.m00 * ox + g.m01 * oy;
      const double y = dec_c + g.m10 * ox + g.m11 * oy;
      // this ray's slot: dS/dbeta, written by the first source's launch and read by the inline lens
      // launches; with several planes, beta, written by mp_grad_kernel and read by the source launches
      T *gb = (gbeta && (k < m.n_lens || k == m.n_defl || (m.mp_n && k < m.n_defl + m.n_source)))
                  ? gbeta + ((size_t(img) * npix + pix) * size_t(g.s * g.s) + size_t(sy * g.s + sx)) * 2
                  : nullptr;
      if (gb) CHECK_SPAN(gb, 2 * sizeof(T));
      if constexpr (type == LENSIM_EPL) lens_ray_grad<T, HALO>(m, ga, rp, img, k, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_SIE) lens_ray_grad<T, HALO, LENSIM_SIE>(m, ga, rp, img, k, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_NIE) lens_ray_grad<T, HALO, LENSIM_NIE>(m, ga, rp, img, k, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_NFW) halo_ray_grad<T, 4, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_TNFW) halo_ray_grad<T, 5, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_SHEAR) ray_grad<T, 4, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_CONVERGENCE)
        ray_grad<T, 3, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_SERSIC) ray_grad<T, 5, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_SERSIC_ELLIPSE)
        ray_grad<T, 7, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
      else if constexpr (type == LENSIM_GAUSSIAN) ray_grad<T, 4, HALO>(m, ga, rp, img, k, type, x, y, w, acc, gb);
    }
  }
  }
  // block sum of the 8 accumulators in a fixed order
  __shared__ double red[256 / 64][LENSIM_NPARAM];
#pragma unroll
  for (int q = 0; q < LENSIM_NPARAM; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < LENSIM_NPARAM) {
    double v = 0.0;
    for (int wv = 0; wv < int(blockDim.x >> 6); ++wv) v += red[wv][threadIdx.x];
    AT(ga.gpart, ((size_t(img) * m.n_prof + k) * gridDim.x + blockIdx.x) * LENSIM_NPARAM + threadIdx.x) = v;
  }
}

// grad_kernel for profile k of the model (its type read on the host).
template <typename T, bool HALO>
int launch_grad_profile(const ModelArgs &m, const GridArgs &g, const GradArgs &ga, const double *recd,
                        const float *recf, int nblk, int n_img, int k, hipStream_t st) {
  const dim3 grid(nblk, n_img);
#define LENSIM_GRAD(TY)                                                                                     \
  grad_kernel<T, HALO, TY><<<grid, 256, 0, st>>>(m, g, ga, recd, recf, ga.coefd, ga.coeff, k)
  switch (type_of(m, k)) {
    case LENSIM_EPL: LENSIM_GRAD(LENSIM_EPL); break;
    case LENSIM_SIE: LENSIM_GRAD(LENSIM_SIE); break;
    case LENSIM_NIE: LENSIM_GRAD(LENSIM_NIE); break;
    case LENSIM_NFW:  // present only in a model with halos (HALO)
      if constexpr (HALO) LENSIM_GRAD(LENSIM_NFW); else LENSIM_GRAD(LENSIM_NULL);
      break;
    case LENSIM_TNFW:
      if constexpr (HALO) LENSIM_GRAD(LENSIM_TNFW); else LENSIM_GRAD(LENSIM_NULL);
      break;
    case LENSIM_SHEAR: LENSIM_GRAD(LENSIM_SHEAR); break;
    case LENSIM_CONVERGENCE: LENSIM_GRAD(LENSIM_CONVERGENCE); break;
    case LENSIM_SERSIC: LENSIM_GRAD(LENSIM_SERSIC); break;
    case LENSIM_SERSIC_ELLIPSE: LENSIM_GRAD(LENSIM_SERSIC_ELLIPSE); break;
    case LENSIM_GAUSSIAN: LENSIM_GRAD(LENSIM_GAUSSIAN); break;
    default: LENSIM_GRAD(LENSIM_NULL); break;  // no kwargs: the kernel writes zero partials
  }
#undef LENSIM_GRAD
  return check_launch("grad_kernel");
}

// d flux / d(kwargs) of the bulk deflectors (LensModelBulk lists, records
// n_lens .. n_defl - 1).  With beta = theta - alpha_scaling sum_j alpha_j,
// d(ray flux)/d p_k = -alpha_scaling grad S(beta) . d alpha_k / d p_k, so a
// ray's total deflection and source gradient are needed once, and then each
// deflector's own deflection in duals.  Block (b, img, c): the rays of pixels
// [b kGradPix 256, (b + 1) kGradPix 256) of image img and deflectors
// n_lens + c dpb + d, d < dpb (dpb = 64, 128 or 256 threads per deflector
// group, 256 / dpb groups of whole waves).  Per tile of 256 rays each thread
// traces one ray into LDS: (x, y) and g = w (-alpha_scaling) dS/dbeta; then
// group grp's thread d runs its deflector's duals over the tile's rays grp,
// grp + 256/dpb, ... (the same ray for every lane of a wave: the zero-weight
// skip is uniform), and the groups' sums are added in a fixed order.
// TYPE is the list's common type, or LENSIM_NULL for a list of several types
// (a switch per deflector, outside the ray loop).
// The deflectors' duals run in double in both forms: the NFW family's closed
// forms cancel inside and outside the scale radius (the float forward
// rewrites them; their duals would not follow), which costs a float dual all
// its precision in the derivative by the centre.
template <int TYPE, int N>
__device__ __forceinline__ void bulk_tile_duals(const double *pk, const double *rd, const double *lx,
                                                const double *ly, const double *lgx, const double *lgy, int r0,
                                                int step, int r1, double *acc) {
  if constexpr (TYPE == LENSIM_SIE || TYPE == LENSIM_NIE) {  // orientation from the record (nie_dalpha)
    for (int r = r0; r < r1; r += step) {
      const double gx = lgx[r], gy = lgy[r];
      if (gx == 0.0 && gy == 0.0) continue;
      double a_x, a_y, dax[6], day[6];
      nie_dalpha<double, TYPE>(rd, pk, lx[r], ly[r], a_x, a_y, dax, day);
#pragma unroll
      for (int i = 0; i < N; ++i) acc[i] += gx * dax[i] + gy * day[i];
    }
    return;
  }
  using S = gm::Dual<double, N>;
  S ps[LENSIM_NPARAM];
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) ps[i] = i < N ? S::seed(pk[i], i) : S(pk[i]);
  for (int r = r0; r < r1; r += step) {
    const double gx = lgx[r], gy = lgy[r];
    if (gx == 0.0 && gy == 0.0) continue;
    S ax = S(0.0), ay = S(0.0);
    gm::defl_g<double, S>(TYPE, ps, static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), 1,
                          S(lx[r]), S(ly[r]), ax, ay);
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] += gx * ax.d[i] + gy * ay.d[i];
  }
}

template <typename T, bool HALO, int TYPE>
__global__ __launch_bounds__(256) void bulk_grad_kernel(ModelArgs m, GridArgs g, GradArgs ga,
                                                        const double *__restrict__ recd,
                                                        const float *__restrict__ recf,
                                                        const double *__restrict__ coefd,
                                                        const float *__restrict__ coeff, int dpb) {
  const int img = blockIdx.y;
  const int npix = g.nx * g.ny, ss = g.s * g.s;
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  const int t = threadIdx.x, d = t % dpb, grp = t / dpb, ngrp = blockDim.x / dpb;
  const int k = m.n_lens + blockIdx.z * dpb + d;
  const bool mine = k < m.n_defl;
  int type = LENSIM_NULL;
  if (mine) type = TYPE != LENSIM_NULL ? TYPE : AT(m.bulk_types, k - m.n_lens);
  double pk[LENSIM_NPARAM], acc[LENSIM_NPARAM];
#pragma unroll
  for (int q = 0; q < LENSIM_NPARAM; ++q) {
    pk[q] = mine ? AT(ga.params, (size_t(img) * m.n_prof + k) * LENSIM_NPARAM + q) : 0.0;
    acc[q] = 0.0;
  }
  const double *rdk = rp.recd + (mine ? k : 0) * kRec;  // this deflector's record (SIE / NIE orientation)
  __shared__ double lx[256], ly[256], lgx[256], lgy[256];
  __shared__ double red[256][LENSIM_NPARAM];
  const int p0 = blockIdx.x * kGradPix * 256, p1 = min(npix, p0 + kGradPix * 256);
  const int r_end = (p1 - p0) * ss;  // rays of the block, pixel-major, s x s sub-rays per pixel
  const double inv_s = g.inv_s, o0 = g.o0, as = m.alpha_scaling;
  for (int base = 0; base < r_end; base += 256) {
    // one ray per thread: position, weight and the source gradient at its beta
    {
      const int rr = base + t;
      const bool active = rr < r_end;
      int sub0;
      const int q = div_by(active ? rr : 0, ss, g.inv_ss, sub0);
      const int pix = p0 + q, sub = sub0;
      int i, sx;
      const int j = div_by(pix, g.nx, g.inv_nx, i), sy = div_by(sub, g.s, g.inv_s, sx);
      const double ox = o0 + sx * inv_s, oy = o0 + sy * inv_s;
      const double x = i * g.m00 + j * g.m01 + g.ra0 + g.m00 * ox + g.m01 * oy;
      const double y = i * g.m10 + j * g.m11 + g.dec0 + g.m10 * ox + g.m11 * oy;
      const double w = active ? double(AT(static_cast<const T *>(ga.glow), size_t(img) * npix + pix)) * g.scale *
                                    m.flux_scaling
                              : 0.0;
      double gx = 0.0, gy = 0.0;
      if (w != 0.0 && m.n_source > 0 && evaluated(g, i, j)) {
        double ax = 0.0, ay = 0.0;
        deflect_all<T, HALO>(m, rp, img, x, y, ax, ay);
        const gm::Dual<T, 2> f2 = sources_at<T>(m, ga, img, T(x - ax * as), T(y - ay * as));
        gx = -as * w * double(f2.d[0]);
        gy = -as * w * double(f2.d[1]);
      }
      lx[t] = x;
      ly[t] = y;
      lgx[t] = gx;
      lgy[t] = gy;
    }
    __syncthreads();
    const int r1 = min(256, r_end - base);
    if constexpr (TYPE != LENSIM_NULL) {
      if (mine) bulk_tile_duals<TYPE, n_params_of(TYPE)>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc);
    } else {
      switch (type) {
        case LENSIM_SIE: bulk_tile_duals<LENSIM_SIE, 5>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc); break;
        case LENSIM_NIE: bulk_tile_duals<LENSIM_NIE, 6>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc); break;
        case LENSIM_NFW:
          if constexpr (HALO) bulk_tile_duals<LENSIM_NFW, 4>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc);
          break;
        case LENSIM_TNFW:
          if constexpr (HALO) bulk_tile_duals<LENSIM_TNFW, 5>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc);
          break;
        case LENSIM_SHEAR: bulk_tile_duals<LENSIM_SHEAR, 4>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc); break;
        case LENSIM_CONVERGENCE:
          bulk_tile_duals<LENSIM_CONVERGENCE, 3>(pk, rdk, lx, ly, lgx, lgy, grp, ngrp, r1, acc);
          break;
        default: break;  // no kwargs: zero partials
      }
    }
    __syncthreads();
  }
  // the groups' sums in group order
#pragma unroll
  for (int q = 0; q < LENSIM_NPARAM; ++q) red[t][q] = acc[q];
  __syncthreads();
  if (grp == 0 && mine) {
#pragma unroll
    for (int q = 0; q < LENSIM_NPARAM; ++q) {
      double v = 0.0;
      for (int gi = 0; gi < ngrp; ++gi) v += red[gi * dpb + d][q];
      AT(ga.gpart, ((size_t(img) * m.n_prof + k) * gridDim.x + blockIdx.x) * LENSIM_NPARAM + q) = v;
    }
  }
}

template <typename T, bool HALO>
int launch_bulk_grad(const ModelArgs &m, const GridArgs &g, const GradArgs &ga, const double *recd,
                     const float *recf, int nblk, int n_img, hipStream_t st) {
  const int dpb = m.n_bulk <= 64 ? 64 : (m.n_bulk <= 128 ? 128 : 256);
  const dim3 grid(nblk, n_img, (m.n_bulk + dpb - 1) / dpb);
#define LENSIM_BULK_GRAD(TY) \
  bulk_grad_kernel<T, HALO, TY><<<grid, 256, 0, st>>>(m, g, ga, recd, recf, ga.coefd, ga.coeff, dpb)
  switch (m.bulk_type) {
    case LENSIM_SIE: LENSIM_BULK_GRAD(LENSIM_SIE); break;
    case LENSIM_NIE: LENSIM_BULK_GRAD(LENSIM_NIE); break;
    case LENSIM_NFW:
      if constexpr (HALO) LENSIM_BULK_GRAD(LENSIM_NFW);
      break;
    case LENSIM_TNFW:
      if constexpr (HALO) LENSIM_BULK_GRAD(LENSIM_TNFW);
      break;
    case LENSIM_SHEAR: LENSIM_BULK_GRAD(LENSIM_SHEAR); break;
    case LENSIM_CONVERGENCE: LENSIM_BULK_GRAD(LENSIM_CONVERGENCE); break;
    default: LENSIM_BULK_GRAD(LENSIM_NULL); break;  // several types: d_bulk_types per deflector
  }
#undef LENSIM_BULK_GRAD
  return check_launch("bulk_grad_kernel");
}

// Deflector k's deflection at (x, y) and its derivatives by its kwargs,
// slots 0 .. N-1 (N returned; 0 for a type without kwargs): the EPL's closed
// forms in the compute type, the other types' duals in double (as in the bulk
// list's gradient).  The last two slots of every deflector type are its
// centre, whose derivative is minus the column of d alpha / d(x, y).
template <int TYPE, int N>
__device__ __forceinline__ void dalpha_duals(const double *pk, double x, double y, double &ax, double &ay,
                                             double *dax, double *day, double *dcen) {
  using S = gm::Dual<double, N>;
  S ps[LENSIM_NPARAM];
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) ps[i] = i < N ? S::seed(pk[i], i) : S(pk[i]);
  S sx = S(0.0), sy = S(0.0);
  gm::defl_g<double, S>(TYPE, ps, static_cast<const double *>(nullptr), static_cast<const double *>(nullptr), 1,
                        S(x), S(y), sx, sy);
  ax = sx.v;
  ay = sy.v;
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) {
    dax[i] = i < N ? sx.d[i] : 0.0;
    day[i] = i < N ? sy.d[i] : 0.0;
  }
  // the centre's derivatives (slots N-2, N-1): minus d alpha / dx and d alpha / dy
  dcen[0] = sx.d[N - 2];
  dcen[1] = sy.d[N - 2];
  dcen[2] = sx.d[N - 1];
  dcen[3] = sy.d[N - 1];
}

// nie_dalpha in double, laid out as dalpha_duals'.
template <int TYPE>
__device__ __forceinline__ void nie_dalpha_mp(const double *rd, const double *pk, double x, double y, double &ax,
                                              double &ay, double *dax, double *day, double *dcen) {
  constexpr int N = n_params_of(TYPE);
  double dx[6], dy[6];
  nie_dalpha<double, TYPE>(rd, pk, x, y, ax, ay, dx, dy);
#pragma unroll
  for (int i = 0; i < LENSIM_NPARAM; ++i) {
    dax[i] = i < N ? dx[i] : 0.0;
    day[i] = i < N ? dy[i] : 0.0;
  }
  dcen[0] = dx[N - 2];
  dcen[1] = dy[N - 2];
  dcen[2] = dx[N - 1];
  dcen[3] = dy[N - 1];
}

template <typename T, bool HALO>
__device__ __forceinline__ void deflector_dalpha(const ModelArgs &m, const GradArgs &ga, const RecPtrs<T> &rp,
                                                 int img, int k, int type, double x, double y, double &ax,
                                                 double &ay, double *dax, double *day, double *dcen) {
  const double *pk = ga.params + (size_t(img) * m.n_prof + k) * LENSIM_NPARAM;
  CHECK_SPAN(pk, LENSIM_NPARAM * sizeof(double));
  switch (type) {
    case LENSIM_EPL: {
      T a_x, a_y, dx[6], dy[6];
      epl_dalpha<T>(m, ga, rp, img, k, pk, x, y, a_x, a_y, dx, dy);
      ax = double(a_x);
      ay = double(a_y);
#pragma unroll
      for (int i = 0; i < LENSIM_NPARAM; ++i) {
        dax[i] = i < 6 ? double(dx[i]) : 0.0;
        day[i] = i < 6 ? double(dy[i]) : 0.0;
      }
      dcen[0] = dax[4];
      dcen[1] = day[4];
      dcen[2] = dax[5];
      dcen[3] = day[5];
      break;
    }
    case LENSIM_SIE: nie_dalpha_mp<LENSIM_SIE>(rp.recd + k * kRec, pk, x, y, ax, ay, dax, day, dcen); break;
    case LENSIM_NIE: nie_dalpha_mp<LENSIM_NIE>(rp.recd + k * kRec, pk, x, y, ax, ay, dax, day, dcen); break;
    case LENSIM_NFW:
      if constexpr (HALO) dalpha_duals<LENSIM_NFW, 4>(pk, x, y, ax, ay, dax, day, dcen);
      break;
    case LENSIM_TNFW:
      if constexpr (HALO) dalpha_duals<LENSIM_TNFW, 5>(pk, x, y, ax, ay, dax, day, dcen);
      break;
    case LENSIM_SHEAR: dalpha_duals<LENSIM_SHEAR, 4>(pk, x, y, ax, ay, dax, day, dcen); break;
    case LENSIM_CONVERGENCE: dalpha_duals<LENSIM_CONVERGENCE, 3>(pk, x, y, ax, ay, dax, day, dcen); break;
    default: break;
  }
}

// d flux / d(kwargs) of every deflector of a multi-plane model, the
// recursion of trace_ray (MultiPlaneBulk.ray_shooting,
// LensModel/MultiPlane/multi_plane_bulk.py:152-212)
//   P_p = P_{p-1} + T_p A_{p-1},  A_p = A_{p-1} - f_p alpha_p(P_p / Tz_p),
//   beta = P_n / Tz_n  (P_{-1} = 0, A_{-1} = the ray's angle)
// differentiated in reverse.  From the adjoints (gP, gA) of (P_p, A_p):
//   lambda = -f_p gA  (the adjoint of alpha_p),
//   d flux / d kwargs_p = lambda . d alpha_p / d kwargs_p,
//   gP += (d alpha_p / d x)^T lambda / Tz_p,  then gA += T_p gP,
// starting from gP = w dS/dbeta / Tz_n, gA = 0 at the source plane.  The
// plane states are not stored: each backward step inverts its forward one,
// A_{p-1} = A_p + f_p alpha_p and P_{p-1} = P_p - T_p A_{p-1}, with alpha_p
// the value of the duals the step evaluates anyway.  (d alpha / d x)^T lambda
// is minus lambda's product with the centre's derivative.
// Block (b, img) takes sweeps [b spb, (b + 1) spb) of 256 rays of image img
// (the rays pixel-major, s x s sub-rays per pixel), one ray per thread per
// sweep; all threads sweep the planes together, so a plane's contributions
// are summed over the block (waves, then the waves in order) and added to its
// deflector's partial in gpart_mp, which the block owns; mp_partial_reduce
// sums the blocks in order.  Each ray's beta is also stored (ga.gbeta) for
// the source profiles' launches.
template <typename T, bool HALO>
__global__ __launch_bounds__(256) void mp_grad_kernel(ModelArgs m, GridArgs g, GradArgs ga,
                                                      const double *__restrict__ recd,
                                                      const float *__restrict__ recf,
                                                      const double *__restrict__ coefd,
                                                      const float *__restrict__ coeff, double *__restrict__ gpart_mp,
                                                      int spb) {
  const int img = blockIdx.y;
  const int npix = g.nx * g.ny, ss = g.s * g.s;
  const RecPtrs<T> rp = rec_ptrs<T>(m, recd, recf, coefd, coeff, img);
  const int t = threadIdx.x;
  const size_t o = size_t(img) * size_t(m.mp_stride);
  const int nd = m.mp_n - 1;  // deflector planes; plane nd is the source's
  auto slot = [&](int k) {
    return ((size_t(img) * m.n_defl + k) * gridDim.x + blockIdx.x) * LENSIM_NPARAM;
  };
  // this block's partials of every deflector start at zero
  for (int q = t; q < m.n_defl * LENSIM_NPARAM; q += blockDim.x)
    AT(gpart_mp, slot(q / LENSIM_NPARAM) + q % LENSIM_NPARAM) = 0.0;
  __syncthreads();
  __shared__ double red[2][256 / 64][LENSIM_NPARAM];
  const int r_all = npix * ss;
  const int r0 = min(r_all, int(blockIdx.x) * spb * 256), r_end = min(r_all, r0 + spb * 256);
  int step = 0;  // plane steps so far: consecutive ones alternate between the two halves of red
  for (int base = r0; base < r_end; base += 256) {
    const int rr = base + t;
    const bool active = rr < r_end;
    int sub;
    const int pix = div_by(active ? rr : r0, ss, g.inv_ss, sub);
    int i, sx;
    const int j = div_by(pix, g.nx, g.inv_nx, i), sy = div_by(sub, g.s, g.inv_s, sx);
    const double ox = g.o0 + sx * g.inv_s, oy = g.o0 + sy * g.inv_s;
    const double x = i * g.m00 + j * g.m01 + g.ra0 + g.m00 * ox + g.m01 * oy;
    const double y = i * g.m10 + j * g.m11 + g.dec0 + g.m10 * ox + g.m11 * oy;
    const double w = active ? double(AT(static_cast<const T *>(ga.glow), size_t(img) * npix + pix)) * g.scale *
                                  m.flux_scaling
                            : 0.0;
    const bool live = w != 0.0 && m.n_source > 0 && evaluated(g, i, j);
    // forwards to the source plane
    double Ax = x, Ay = y, Px = 0.0, Py = 0.0;
    double gPx = 0.0, gPy = 0.0, gAx = 0.0, gAy = 0.0;
    if (live) {
      for (int p = 0; p <= nd; ++p) {
        const double dt = AT(m.mp_tij, o + p);
        Px += Ax * dt;
        Py += Ay * dt;
        if (p < nd) {
          const double tz = AT(m.mp_tz, o + p);
          double ax = 0.0, ay = 0.0;
          int k = AT(m.mp_order, p);
          k = k < 0 ? 0 : (k >= m.n_defl ? m.n_defl - 1 : k);
          deflect_one<T, HALO, true>(m, rp, img, k, Px / tz, Py / tz, ax, ay);
          const double f = AT(m.mp_r2p, o + p);
          Ax -= ax * f;
          Ay -= ay * f;
        }
      }
      const double tzn = AT(m.mp_tz, o + nd);
      const T bx = T(Px / tzn), by = T(Py / tzn);
      if (ga.gbeta) {  // beta for the source launches
        T *gb = static_cast<T *>(ga.gbeta) + ((size_t(img) * npix + pix) * size_t(ss) + size_t(sub)) * 2;
        AT(gb, 0) = bx;
        AT(gb, 1) = by;
      }
      const gm::Dual<T, 2> f2 = sources_at<T>(m, ga, img, bx, by);
      gPx = w * double(f2.d[0]) / tzn;
      gPy = w * double(f2.d[1]) / tzn;
      // the source plane's step (no deflector): A_{nd-1} = A_nd
      const double dt = AT(m.mp_tij, o + nd);
      gAx += dt * gPx;
      gAy += dt * gPy;
      Px -= dt * Ax;
      Py -= dt * Ay;
    }
    // backwards over the deflector planes
    for (int p = nd - 1; p >= 0; --p) {
      int k = AT(m.mp_order, p);
      k = k < 0 ? 0 : (k >= m.n_defl ? m.n_defl - 1 : k);
      const int type = k < m.n_lens ? m.types[k] : (m.bulk_type ? m.bulk_type : AT(m.bulk_types, k - m.n_lens));
      const int n = (type == LENSIM_EPL && k >= m.n_lens) ? 0 : n_params_of(type);  // bulk EPL: NULL
      double c[LENSIM_NPARAM];
#pragma unroll
      for (int q = 0; q < LENSIM_NPARAM; ++q) c[q] = 0.0;
      if (live && n > 0) {
        const double tz = AT(m.mp_tz, o + p), f = AT(m.mp_r2p, o + p), dt = AT(m.mp_tij, o + p);
        double ax = 0.0, ay = 0.0, dax[LENSIM_NPARAM], day[LENSIM_NPARAM], dcen[4];
        deflector_dalpha<T, HALO>(m, ga, rp, img, k, type, Px / tz, Py / tz, ax, ay, dax, day, dcen);
        const double lx = -f * gAx, ly = -f * gAy;
#pragma unroll
        for (int q = 0; q < LENSIM_NPARAM; ++q) c[q] = lx * dax[q] + ly * day[q];
        gPx -= (lx * dcen[0] + ly * dcen[1]) / tz;
        gPy -= (lx * dcen[2] + ly * dcen[3]) / tz;
        Ax += f * ax;
        Ay += f * ay;
        gAx += dt * gPx;
        gAy += dt * gPy;
        Px -= dt * Ax;
        Py -= dt * Ay;
      } else if (live) {  // a deflector without kwargs deflects by zero
        const double dt = AT(m.mp_tij, o + p);
        gAx += dt * gPx;
        gAy += dt * gPy;
        Px -= dt * Ax;
        Py -= dt * Ay;
      }
      // the plane's sum over the block, added to deflector k's partial
      const int buf = step++ & 1;
#pragma unroll
      for (int q = 0; q < LENSIM_NPARAM; ++q) {
        const double v = wave_sum(c[q]);
        if ((t & 63) == 0) red[buf][t >> 6][q] = v;
      }
      __syncthreads();
      if (t < n) {
        double v = 0.0;
        for (int wv = 0; wv < int(blockDim.x >> 6); ++wv) v += red[buf][wv][t];
        // the block's first sweep stores (no wait for a load), later ones add
        if (base == r0)
          AT(gpart_mp, slot(k) + t) = v;
        else
          AT(gpart_mp, slot(k) + t) += v;
      }
    }
  }
}

// The deflectors' partials of the nb_mp blocks of mp_grad_kernel summed in
// block order into the first of the nblk slots grad_reduce_kernel reads (the
// others zero).
__global__ void mp_partial_reduce(const double *__restrict__ gpart_mp, int nb_mp, int n_img, int n_defl, int n_prof,
                                  int nblk, double *__restrict__ gpart) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_img * n_defl * LENSIM_NPARAM) return;
  const int q = i % LENSIM_NPARAM, r = i / LENSIM_NPARAM, k = r % n_defl, img = r / n_defl;
  double v = 0.0;
  for (int b = 0; b < nb_mp; ++b) v += AT(gpart_mp, (size_t(r) * nb_mp + b) * LENSIM_NPARAM + q);
  const size_t o = (size_t(img) * n_prof + k) * nblk;
  for (int b = 0; b < nblk; ++b) AT(gpart, (o + b) * LENSIM_NPARAM + q) = b == 0 ? v : 0.0;
}

template <typename T>
int launch_mp_grad(const ModelArgs &m, const GridArgs &g, const GradArgs &ga, const double *recd, const float *recf,
                   double *gpart_mp, int nblk, int n_img, hipStream_t st) {
  const size_t rays = size_t(g.nx) * g.ny * g.s * g.s;
  const int nb_mp = mp_grad_blocks(rays, n_img);
  const int spb = int(((rays + 255) / 256 + nb_mp - 1) / nb_mp);
  const dim3 grid(nb_mp, n_img);
  if (m.has_halo)
    mp_grad_kernel<T, true><<<grid, 256, 0, st>>>(m, g, ga, recd, recf, ga.coefd, ga.coeff, gpart_mp, spb);
  else
    mp_grad_kernel<T, false><<<grid, 256, 0, st>>>(m, g, ga, recd, recf, ga.coefd, ga.coeff, gpart_mp, spb);
  int rc = check_launch("mp_grad_kernel");
  if (rc || m.n_defl == 0) return rc;
  const int n = n_img * m.n_defl * LENSIM_NPARAM;
  mp_partial_reduce<<<(n + 255) / 256, 256, 0, st>>>(gpart_mp, nb_mp, n_img, m.n_defl, m.n_prof, nblk, ga.gpart);
  return check_launch("mp_partial_reduce");
}

__global__ void grad_reduce_kernel(const double *__restrict__ gpart, int nblk, int n, double *__restrict__ grad) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (img * n_prof + k) * 8 + q
  if (idx >= n) return;
  const int q = idx % LENSIM_NPARAM, ik = idx / LENSIM_NPARAM;
  double s = 0.0;
  for (int b = 0; b < nblk; ++b) s += AT(gpart, (size_t(ik) * nblk + b) * LENSIM_NPARAM + q);
  AT(grad, idx) = s;
}

// One wave per image: lane l sums partials l, l + 64, ... and a fixed
// butterfly adds the 64 lane sums, so the order depends on n_tiles alone (not
// on the batch); the partials of an image are read contiguously.
__global__ __launch_bounds__(64) void logl_kernel(const double *__restrict__ partial, int n_tiles,
                                                  double *__restrict__ logl) {
  const int img = blockIdx.x;
  const double *p = partial + size_t(img) * n_tiles;
  double s = 0.0;
  for (int t = threadIdx.x; t < n_tiles; t += 64) s += AT(p, t);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (threadIdx.x == 0) AT(logl, img) = -s / 2.0;
}

// ------------------------------------------------------------ launches ---

int check_launch(const char *where) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, where);
  return LENSIM_OK;
}

int launch_prep(const ModelArgs &m, const double *params, int n_img, unsigned char *ws, const Layout &L,
                hipStream_t st, bool with_dcoef = false) {
  const int n = n_img * m.n_prof;
  if (n == 0) return LENSIM_OK;
  prep_kernel<<<(n + 63) / 64, 64, 0, st>>>(m, params, n_img, reinterpret_cast<double *>(ws + L.recd),
                                             reinterpret_cast<float *>(ws + L.recf),
                                             reinterpret_cast<double *>(ws + L.coefd),
                                             reinterpret_cast<float *>(ws + L.coeff),
                                             with_dcoef ? reinterpret_cast<double *>(ws + L.dcoefd) : nullptr,
                                             with_dcoef ? reinterpret_cast<float *>(ws + L.dcoeff) : nullptr);
  return check_launch("prep_kernel");
}

int grid_args(const lensim_grid *grid, GridArgs &g) {
  if (!grid) return fail(LENSIM_EINVAL, "grid is NULL");
  if (grid->nx <= 0 || grid->ny <= 0 || grid->supersampling <= 0)
    return fail(LENSIM_EINVAL, "grid sizes must be positive (nx=%d ny=%d s=%d)", grid->nx, grid->ny,
                grid->supersampling);
  if (int64_t(grid->nx) * grid->ny > (int64_t(1) << 30)) return fail(LENSIM_EINVAL, "grid too large");
  g.nx = grid->nx;
  g.ny = grid->ny;
  g.s = grid->supersampling;
  g.ra0 = grid->ra_at_xy_0;
  g.dec0 = grid->dec_at_xy_0;
  g.m00 = grid->pix2angle[0];
  g.m01 = grid->pix2angle[1];
  g.m10 = grid->pix2angle[2];
  g.m11 = grid->pix2angle[3];
  const double det = g.m00 * g.m11 - g.m01 * g.m10;
  const double pixel_width = sqrt(fabs(det));  // lenstronomy PixelGrid.pixel_width
  g.scale = pixel_width * pixel_width / (double(g.s) * double(g.s));
  g.mask = grid->d_eval_mask;
  g.mask_nx = g.nx;
  g.mask_s = 1;
  grid_consts(g);
  return LENSIM_OK;
}

// The model with the sub-ray steps of grid g (SubSteps) for its records.
ModelArgs with_grid(const ModelArgs &m, const GridArgs &g) {
  ModelArgs mg = m;
  const double inv_s = 1.0 / g.s;
  mg.sub.sx[0] = g.m00 * inv_s;
  mg.sub.sx[1] = g.m10 * inv_s;
  mg.sub.rr[0] = g.m01 * inv_s - g.s * mg.sub.sx[0];
  mg.sub.rr[1] = g.m11 * inv_s - g.s * mg.sub.sx[1];
  mg.sub.asc = m.alpha_scaling;
  return mg;
}

int psf_dims(const lensim_psf *psf, uint32_t flags, int &kx, int &ky, int &edge) {
  const bool none = !psf || psf->type == LENSIM_PSF_NONE || (flags & LENSIM_IMAGE_UNCONVOLVED);
  if (none) {
    kx = ky = 1;
    edge = 0;
    return LENSIM_OK;
  }
  if (psf->type != LENSIM_PSF_PIXEL && psf->type != LENSIM_PSF_GAUSSIAN)
    return fail(LENSIM_EUNSUPPORTED, "psf type %d not supported", psf->type);
  if (psf->kx <= 0 || psf->ky <= 0 || !(psf->kx & 1) || !(psf->ky & 1))
    return fail(LENSIM_EINVAL, "psf kernel must have odd positive sizes (got %d x %d)", psf->ky, psf->kx);
  if (!psf->d_kernel) return fail(LENSIM_EINVAL, "psf kernel pointer is NULL");
  kx = psf->kx;
  ky = psf->ky;
  edge = psf->type == LENSIM_PSF_GAUSSIAN;
  return LENSIM_OK;
}

// Supersampled-convolution plan of a PSF on grid g (off when unconvolved,
// without a PSF, or when the grid is not supersampled).
int psf_ss(const lensim_psf *psf, const GridArgs &g, uint32_t flags, SSPlan &ss) {
  ss = SSPlan();
  const bool none = !psf || psf->type == LENSIM_PSF_NONE || (flags & LENSIM_IMAGE_UNCONVOLVED);
  if (none || !psf->supersampled || g.s == 1) return LENSIM_OK;
  if (psf->supersampled != 1) return fail(LENSIM_EINVAL, "psf.supersampled must be 0 or 1");
  if (int64_t(g.nx) * g.ny * g.s * g.s > (int64_t(1) << 30)) return fail(LENSIM_EINVAL, "supersampled grid too large");
  ss.on = 1;
  ss.s = g.s;
  if (psf->kx_low || psf->ky_low || psf->d_kernel_low) {
    if (psf->type != LENSIM_PSF_PIXEL) return fail(LENSIM_EINVAL, "a split kernel needs a PIXEL psf");
    if (psf->kx_low <= 0 || psf->ky_low <= 0 || !(psf->kx_low & 1) || !(psf->ky_low & 1))
      return fail(LENSIM_EINVAL, "low-resolution kernel must have odd positive sizes (got %d x %d)", psf->ky_low,
                  psf->kx_low);
    if (!psf->d_kernel_low) return fail(LENSIM_EINVAL, "low-resolution kernel pointer is NULL");
    ss.kxl = psf->kx_low;
    ss.kyl = psf->ky_low;
  }
  return LENSIM_OK;
}

template <typename T>
size_t conv_smem(int kx, int kyp, int ty = ConvTile<T>::TY) {
  return size_t((ty + kyp - 1) | 1) * (ConvTile<T>::TX + kx - 1) * sizeof(T);
}

// conv_kernel over n_img images (c.n_tiles_x is set here, the tile count goes
// to *n_tiles): 64-row tiles where their LDS fits (four waves a block, a
// smaller share of halo rows), 32-row tiles otherwise.
template <typename T, int TY, bool LW>
int launch_conv_tile(const ConvArgs &c, dim3 grid, size_t smem, const T *in, const double *kf, T *out,
                     double *partial, hipStream_t st) {
  if (smem > 65536) {  // dynamic LDS beyond 64 KB needs the kernel attribute raised first
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&conv_kernel<T, TY, LW>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(conv_kernel)");
  }
  conv_kernel<T, TY, LW><<<grid, ConvTile<T>::TX * TY / kWin, smem, st>>>(c, in, kf, out, partial);
  return check_launch("conv_kernel");
}

template <typename T>
int launch_conv_kernel(ConvArgs c, int n_img, const T *in, const double *kf, T *out, double *partial,
                       int *n_tiles, hipStream_t st) {
  constexpr int TX = ConvTile<T>::TX;
  int ty = ConvTile<T>::TY;
  if (conv_smem<T>(c.kx, c.kyp, 64) <= kMaxLds) ty = 64;
  c.n_tiles_x = (c.nx + TX - 1) / TX;
  *n_tiles = c.n_tiles_x * ((c.ny + ty - 1) / ty);
  const size_t smem = conv_smem<T>(c.kx, c.kyp, ty);
  if (smem > kMaxLds) return fail(LENSIM_EUNSUPPORTED, "psf %dx%d needs %zu B of LDS per tile", c.ky, c.kx, smem);
  const dim3 grid(*n_tiles, n_img);
  // weights in LDS when the launch has fewer blocks than the GPU has CUs
  const size_t smem_lw = ((smem + 7) & ~size_t(7)) + size_t(c.kx) * c.kyp * sizeof(double);
  const bool lw = size_t(*n_tiles) * n_img < 256 && smem_lw <= kMaxLds;
  if (ty == 64)
    return lw ? launch_conv_tile<T, 64, true>(c, grid, smem_lw, in, kf, out, partial, st)
              : launch_conv_tile<T, 64, false>(c, grid, smem, in, kf, out, partial, st);
  return lw ? launch_conv_tile<T, ConvTile<T>::TY, true>(c, grid, smem_lw, in, kf, out, partial, st)
            : launch_conv_tile<T, ConvTile<T>::TY, false>(c, grid, smem, in, kf, out, partial, st);
}

// Whether the mixed form convolves on the matrix cores (any real PSF whose
// MFMA tile fits in 64 KB of LDS) or on the vector ALUs.
template <typename T>
bool use_mfma(bool identity, int kx, int ky) {
  if (sizeof(T) != sizeof(float) || identity) return false;
  return mfma_nks(kx) <= kMfMaxNks && mfma_smem(kx, ky, 2) <= kMaxLds;
}

template <int NKS, int TL>
int launch_conv_mfma_n(const ConvArgs &c, dim3 grid, const float *in, const int *hdr, const half8 *frag,
                       const float *blkmax, int nblk, float *out, double *partial, hipStream_t st) {
  const size_t smem = mfma_smem(c.kx, c.ky, TL);
  if (smem > 65536) {  // dynamic LDS beyond 64 KB needs the kernel attribute raised first
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&conv_mfma_kernel<NKS, TL>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(conv_mfma_kernel)");
  }
  conv_mfma_kernel<NKS, TL><<<grid, 256, smem, st>>>(c, mfma_ls(c.kx, TL), in, hdr, frag, blkmax, nblk, out,
                                                      partial);
  return check_launch("conv_mfma_kernel");
}

template <int NKS>
int launch_conv_mfma_tl(const ConvArgs &c, int tl, dim3 grid, const float *in, const int *hdr, const half8 *frag,
                        const float *blkmax, int nblk, float *out, double *partial, hipStream_t st) {
  switch (tl) {
    case 2: return launch_conv_mfma_n<NKS, 2>(c, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    case 3: return launch_conv_mfma_n<NKS, 3>(c, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    case 4: return launch_conv_mfma_n<NKS, 4>(c, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    default: return fail(LENSIM_EINVAL, "matrix-core tile width %d", tl);
  }
}

// Whether the FP64 form convolves on the matrix cores (kernels of at least
// 21 x 21 whose tile fits the LDS; smaller ones use the direct stencil).
template <typename T>
bool use_mf64(bool identity, int kx, int ky) {
  if (sizeof(T) != sizeof(double) || identity) return false;
  return kx >= 21 && ky >= 21 && mf64_smem(kx, ky, 2) <= kMaxLds;
}

template <int TL>
int launch_conv_mf64_n(const ConvArgs &c, dim3 grid, const double *in, const double *frag, double *out,
                       double *partial, hipStream_t st) {
  const size_t smem = mf64_smem(c.kx, c.ky, TL);
  if (smem > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&conv_mf64_kernel<TL>),
                                             hipFuncAttributeMaxDynamicSharedMemorySize, int(smem));
    if (e != hipSuccess) return hip_fail(e, "hipFuncSetAttribute(conv_mf64_kernel)");
  }
  conv_mf64_kernel<TL><<<grid, 256, smem, st>>>(c, mf64_ls(c.kx, TL), mf64_nk4(c.kx), in, frag, out, partial);
  return check_launch("conv_mf64_kernel");
}

int launch_conv_mf64(ConvArgs c, const MfmaPlan &pl, int n_img, const double *in, const double *frag, double *out,
                     double *partial, hipStream_t st) {
  if (pl.tl == 0) return fail(LENSIM_EUNSUPPORTED, "psf %dx%d exceeds the float64 matrix-core tile's LDS", c.ky, c.kx);
  c.n_tiles_x = pl.n_tiles_x;
  const dim3 grid(pl.n_tiles, n_img);
  switch (pl.tl) {
    case 2: return launch_conv_mf64_n<2>(c, grid, in, frag, out, partial, st);
    case 3: return launch_conv_mf64_n<3>(c, grid, in, frag, out, partial, st);
    case 4: return launch_conv_mf64_n<4>(c, grid, in, frag, out, partial, st);
    default: return fail(LENSIM_EINVAL, "matrix-core tile width %d", pl.tl);
  }
}

// The MFMA convolution of n_img images by plan pl (c.kx, c.ky, c.edge set;
// c.n_tiles_x is set here).
int launch_conv_mfma(ConvArgs c, const MfmaPlan &pl, int n_img, const float *in, const int *hdr, const half8 *frag,
                     const float *blkmax, int nblk, float *out, double *partial, hipStream_t st) {
  if (pl.tl == 0) return fail(LENSIM_EUNSUPPORTED, "psf %dx%d exceeds the matrix-core tile's LDS", c.ky, c.kx);
  c.n_tiles_x = pl.n_tiles_x;
  const dim3 grid(pl.n_tiles, n_img);
  switch (mfma_nks(c.kx)) {
    case 1: return launch_conv_mfma_tl<1>(c, pl.tl, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    case 2: return launch_conv_mfma_tl<2>(c, pl.tl, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    case 3: return launch_conv_mfma_tl<3>(c, pl.tl, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    case 4: return launch_conv_mfma_tl<4>(c, pl.tl, grid, in, hdr, frag, blkmax, nblk, out, partial, st);
    default: return fail(LENSIM_EUNSUPPORTED, "psf width %d exceeds the matrix-core convolution", c.kx);
  }
}

// The image pipeline shared by lensim_image and lensim_loglike.
template <typename T>
int run_image(const ModelArgs &m_in, const GridArgs &g, const lensim_psf *psf, int kx, int ky, int edge,
              const lensim_data *data, const double *params, int n_img, unsigned flags, T *out, double *logl,
              unsigned char *ws, const Layout &L, hipStream_t st, T *gout = nullptr) {
  const ModelArgs m = with_grid(m_in, g);
  const bool identity = !psf || psf->type == LENSIM_PSF_NONE || (flags & LENSIM_IMAGE_UNCONVOLVED);
  const bool mfma = use_mfma<T>(identity, kx, ky);
  const bool mf64 = use_mf64<T>(identity, kx, ky);
  const int kyp = kyp_of(ky);
  double *kf = reinterpret_cast<double *>(ws + L.psf);
  int *hdr = reinterpret_cast<int *>(ws + L.bfrag);
  half8 *frag = reinterpret_cast<half8 *>(ws + L.bfrag + 16);
  double *frag64 = reinterpret_cast<double *>(ws + L.bfrag);
  // records and the PSF preparation in one launch
  PsfJob pj{};
  pj.mode = mfma ? 2 : (mf64 ? 3 : 1);
  pj.k = identity ? nullptr : psf->d_kernel;
  pj.kx = kx;
  pj.ky = ky;
  pj.kyp = kyp;
  pj.kt = mf64 ? frag64 : kf;
  pj.hdr = hdr;
  pj.frag = frag;
  if (mfma) {
    pj.nks = mfma_nks(kx);
    pj.n_steps = mfma_steps(kx, ky);
  }
  if (mf64) {
    pj.nks = mf64_nk4(kx);
    pj.n_steps = mf64_steps(kx, ky);
  }
  const int nb_prep = (n_img * m.n_prof + 255) / 256;
  const bool dc = gout != nullptr;
  prep_psf_kernel<<<nb_prep + psf_job_blocks(pj), 256, 0, st>>>(
      m, params, n_img, reinterpret_cast<double *>(ws + L.recd), reinterpret_cast<float *>(ws + L.recf),
      reinterpret_cast<double *>(ws + L.coefd), reinterpret_cast<float *>(ws + L.coeff),
      dc ? reinterpret_cast<double *>(ws + L.dcoefd) : nullptr, dc ? reinterpret_cast<float *>(ws + L.dcoeff) : nullptr,
      pj, nb_prep);
  int rc = check_launch("prep_psf_kernel");
  if (rc) return rc;
  T *lowres = reinterpret_cast<T *>(ws + L.lowres);
  float *blkmax = mfma ? reinterpret_cast<float *>(ws + L.blkmax) : nullptr;
  const int npix = g.nx * g.ny;
  timing_mark(0, st);
  const double *recd = reinterpret_cast<const double *>(ws + L.recd);
  const float *recf = reinterpret_cast<const float *>(ws + L.recf);
  const double *coefd = reinterpret_cast<const double *>(ws + L.coefd);
  const float *coeff = reinterpret_cast<const float *>(ws + L.coeff);
  rc = launch_render<T>(m, g, recd, recf, coefd, coeff, flags, lowres, blkmax, n_img, st);
  if (rc) return rc;
  timing_mark(1, st);
  ConvArgs c{};
  c.nx = g.nx;
  c.ny = g.ny;
  c.kx = kx;
  c.ky = ky;
  c.kyp = kyp;
  c.edge = edge;
  double *partial = nullptr;
  if (data) {
    c.data = data->d_data;
    c.noise = data->d_noise_map;
    c.expo = data->d_exposure_map;
    c.mask = data->d_mask;
    c.exposure_time = data->exposure_time;
    c.bkg2 = data->background_rms * data->background_rms;
    c.inv_t = 1.0 / data->exposure_time;
    c.gout = gout;
    c.dstride = data->image_stride;
    partial = reinterpret_cast<double *>(ws + L.partial);
  }
  int n_tiles;
  if (mfma) {
    const MfmaPlan pl = mfma_plan(g.nx, g.ny, n_img, kx, ky);
    n_tiles = n16_tiles(g.nx, g.ny);  // chi^2 partials of the 16 x 16 output tiles
    rc = launch_conv_mfma(c, pl, n_img, reinterpret_cast<const float *>(lowres), hdr, frag, blkmax,
                          (npix + 255) / 256, reinterpret_cast<float *>(out), partial, st);
  } else if (mf64) {
    const MfmaPlan pl = mfma_plan(g.nx, g.ny, n_img, kx, ky, true);
    n_tiles = n16_tiles(g.nx, g.ny);
    rc = launch_conv_mf64(c, pl, n_img, reinterpret_cast<const double *>(lowres), frag64,
                          reinterpret_cast<double *>(out), partial, st);
  } else {
    rc = launch_conv_kernel<T>(c, n_img, lowres, kf, out, partial, &n_tiles, st);
  }
  if (rc) return rc;
  timing_mark(2, st);
  if (data) {
    logl_kernel<<<n_img, 64, 0, st>>>(partial, n_tiles, logl);
    rc = check_launch("logl_kernel");
    if (rc) return rc;
  }
  return LENSIM_OK;
}

// s x s block means of [n_img][ny s][nx s] into [n_img][ny][nx]
// (Util/image_util.py re_size).
template <typename T>
__global__ void pool_kernel(const T *__restrict__ in, int nx, int ny, int s, T *__restrict__ out) {
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  if (pix >= nx * ny) return;
  const int img = blockIdx.y;
  const int y = pix / nx, x = pix - y * nx;
  const int nxh = nx * s;
  const T *src = in + size_t(img) * nx * ny * s * s + size_t(y) * s * nxh + size_t(x) * s;
  double acc = 0.0;
  for (int sy = 0; sy < s; ++sy)
    for (int sx = 0; sx < s; ++sx) acc += double(AT(src, size_t(sy) * nxh + sx));
  AT(out, size_t(img) * nx * ny + pix) = T(acc / (double(s) * double(s)));
}

// Final image of a supersampled convolution: the s x s mean of the convolved
// ray fluxes plus the low-resolution part (if any), with the chi^2 of every
// pixel reduced to one partial per block.
template <typename T>
__global__ __launch_bounds__(256) void combine_kernel(ConvArgs c, const T *__restrict__ chr, const T *__restrict__ lowc,
                                                      int s, T *__restrict__ out, double *__restrict__ partial) {
  const int img = blockIdx.y;
  const int npix = c.nx * c.ny;
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  double chi = 0.0;
  if (pix < npix) {
    const int y = pix / c.nx, x = pix - y * c.nx;
    const int nxh = c.nx * s;
    const T *src = chr + size_t(img) * npix * s * s + size_t(y) * s * nxh + size_t(x) * s;
    double acc = 0.0;
    for (int sy = 0; sy < s; ++sy)
      for (int sx = 0; sx < s; ++sx) acc += double(AT(src, size_t(sy) * nxh + sx));
    double v = acc / (double(s) * double(s));
    if (lowc) v += double(AT(lowc, size_t(img) * npix + pix));
    const T m = T(v);
    if (out) AT(out, size_t(img) * npix + pix) = m;
    if (partial) chi = chi2_pixel<T>(c, pix, m, img);
    if (c.gout) AT(static_cast<T *>(c.gout), size_t(img) * npix + pix) = T(dlogl_dm(c, pix, double(m), img));
  }
  if (partial) {
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = chi;
    __syncthreads();
    if (threadIdx.x == 0) AT(partial, size_t(img) * gridDim.x + blockIdx.x) = red[0] + red[1] + red[2] + red[3];
  }
}

// chi^2 of given model images, one partial per block of 256 pixels.
template <typename T>
__global__ __launch_bounds__(256) void data_chi2_kernel(ConvArgs c, const T *__restrict__ model,
                                                        double *__restrict__ partial) {
  const int img = blockIdx.y;
  const int npix = c.nx * c.ny;
  const int pix = blockIdx.x * blockDim.x + threadIdx.x;
  double chi = 0.0;
  if (pix < npix) chi = chi2_pixel<T>(c, pix, AT(model, size_t(img) * npix + pix), img);
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = chi;
  __syncthreads();
  if (threadIdx.x == 0) AT(partial, size_t(img) * gridDim.x + blockIdx.x) = red[0] + red[1] + red[2] + red[3];
}

// Launches conv_kernel<T> ('same', zero or edge padding) of [n_img][ny][nx]
// with a prepared transposed kernel, no chi^2.
template <typename T>
int launch_conv_plain(int nx, int ny, int kx, int ky, int edge, const T *in, const double *kf, T *out, int n_img,
                      hipStream_t st) {
  ConvArgs c{};
  c.nx = nx;
  c.ny = ny;
  c.kx = kx;
  c.ky = ky;
  c.kyp = kyp_of(ky);
  c.edge = edge;
  int n_tiles = 0;
  return launch_conv_kernel<T>(c, n_img, in, kf, out, nullptr, &n_tiles, st);
}

// The rays of grid g as a grid of their own: pixel (I, J) is sub-ray
// (I mod s, J mod s) of pixel (I / s, J / s) (make_subgrid offsets), with the
// pixel_width^2 scale because the s x s mean follows the convolution.
GridArgs ray_grid(const GridArgs &g) {
  const int s = g.s;
  GridArgs gh = g;
  const double o0 = -0.5 + 1.0 / (2.0 * s);
  gh.nx = g.nx * s;
  gh.ny = g.ny * s;
  gh.s = 1;
  gh.ra0 = g.ra0 + (g.m00 + g.m01) * o0;
  gh.dec0 = g.dec0 + (g.m10 + g.m11) * o0;
  gh.m00 = g.m00 / s;
  gh.m01 = g.m01 / s;
  gh.m10 = g.m10 / s;
  gh.m11 = g.m11 / s;
  gh.scale = g.scale * double(s) * double(s);
  gh.mask_s = g.mask_s * s;
  grid_consts(gh);
  return gh;
}

// Adjoint of the s x s mean: out[J][I] (+)= in[J / s][I / s] / s^2.
template <typename T>
__global__ void unpool_kernel(const T *__restrict__ in, int nx, int ny, int s, int accumulate, T *__restrict__ out) {
  const int nxh = nx * s, nyh = ny * s;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nxh * nyh) return;
  const int img = blockIdx.y;
  const int J = idx / nxh, I = idx - J * nxh;
  const double v = double(AT(in, size_t(img) * nx * ny + size_t(J / s) * nx + I / s)) / (double(s) * double(s));
  T &o = AT(out, size_t(img) * nxh * nyh + idx);
  o = accumulate ? T(double(o) + v) : T(v);
}

// Supersampled convolution (supersampling_convolution=True,
// ImSim/Numerics/numerics.py:103-116, convolution.py:154-166, :238-252): the
// rays are rendered as a (ny s) x (nx s) image (each ray its own pixel),
// convolved there with the supersampled kernel, averaged over s x s, and the
// low-resolution split kernel's convolution of the averaged unconvolved rays
// is added.
template <typename T>
int run_image_ss(const ModelArgs &m_in, const GridArgs &g, const lensim_psf *psf, int kx, int ky, int edge,
                 const SSPlan &ss, const lensim_data *data, const double *params, int n_img, unsigned flags, T *out,
                 double *logl, unsigned char *ws, const Layout &L, hipStream_t st, T *gout = nullptr) {
  const int s = ss.s;
  const GridArgs gh = ray_grid(g);
  const ModelArgs m = with_grid(m_in, gh);
  int rc = launch_prep(m, params, n_img, ws, L, st, gout != nullptr);
  if (rc) return rc;
  const bool mfma = use_mfma<T>(false, kx, ky);
  const int kyp = kyp_of(ky);
  double *kf = reinterpret_cast<double *>(ws + L.psf);
  int *hdr = reinterpret_cast<int *>(ws + L.bfrag);
  half8 *frag = reinterpret_cast<half8 *>(ws + L.bfrag + 16);
  if (mfma) {
    const int n_frag = mfma_steps(kx, ky) * 2 * 64;
    psf_mfma_prep_kernel<<<(n_frag + 255) / 256, 256, 0, st>>>(psf->d_kernel, kx, ky, mfma_nks(kx),
                                                               mfma_steps(kx, ky), hdr, frag);
    rc = check_launch("psf_mfma_prep_kernel");
  } else {
    psf_prep_kernel<<<(kyp * kx + 255) / 256, 256, 0, st>>>(psf->d_kernel, kx, ky, kyp, kf);
    rc = check_launch("psf_prep_kernel");
  }
  if (rc) return rc;
  const int npix = g.nx * g.ny, npix_h = gh.nx * gh.ny;
  T *hr = reinterpret_cast<T *>(ws + L.hr);
  T *chr = reinterpret_cast<T *>(ws + L.chr);
  float *blkmax = mfma ? reinterpret_cast<float *>(ws + L.blkmax) : nullptr;
  const double *recd = reinterpret_cast<const double *>(ws + L.recd);
  const float *recf = reinterpret_cast<const float *>(ws + L.recf);
  const double *coefd = reinterpret_cast<const double *>(ws + L.coefd);
  const float *coeff = reinterpret_cast<const float *>(ws + L.coeff);
  timing_mark(0, st);
  rc = launch_render<T>(m, gh, recd, recf, coefd, coeff, flags, hr, blkmax, n_img, st);
  if (rc) return rc;
  timing_mark(1, st);
  if (mfma) {
    ConvArgs c{};
    c.nx = gh.nx;
    c.ny = gh.ny;
    c.kx = kx;
    c.ky = ky;
    c.kyp = kyp;
    c.edge = edge;
    rc = launch_conv_mfma(c, mfma_plan(gh.nx, gh.ny, n_img, kx, ky), n_img, reinterpret_cast<const float *>(hr),
                          hdr, frag, blkmax, (npix_h + 255) / 256, reinterpret_cast<float *>(chr), nullptr, st);
  } else {
    rc = launch_conv_plain<T>(gh.nx, gh.ny, kx, ky, edge, hr, kf, chr, n_img, st);
  }
  if (rc) return rc;
  T *lowc = nullptr;
  if (ss.kxl > 0) {
    T *lowu = reinterpret_cast<T *>(ws + L.lowu);
    lowc = reinterpret_cast<T *>(ws + L.lowc);
    double *kfl = reinterpret_cast<double *>(ws + L.psf_low);
    pool_kernel<T><<<dim3((npix + 255) / 256, n_img), 256, 0, st>>>(hr, g.nx, g.ny, s, lowu);
    if ((rc = check_launch("pool_kernel"))) return rc;
    const int kypl = kyp_of(ss.kyl);
    psf_prep_kernel<<<(kypl * ss.kxl + 255) / 256, 256, 0, st>>>(psf->d_kernel_low, ss.kxl, ss.kyl, kypl, kfl);
    if ((rc = check_launch("psf_prep_kernel (low)"))) return rc;
    if ((rc = launch_conv_plain<T>(g.nx, g.ny, ss.kxl, ss.kyl, 0, lowu, kfl, lowc, n_img, st))) return rc;
  }
  timing_mark(2, st);
  ConvArgs cc{};
  cc.nx = g.nx;
  cc.ny = g.ny;
  double *partial = nullptr;
  if (data) {
    cc.data = data->d_data;
    cc.noise = data->d_noise_map;
    cc.expo = data->d_exposure_map;
    cc.mask = data->d_mask;
    cc.exposure_time = data->exposure_time;
    cc.bkg2 = data->background_rms * data->background_rms;
    cc.inv_t = 1.0 / data->exposure_time;
    cc.dstride = data->image_stride;
    cc.gout = gout;
    partial = reinterpret_cast<double *>(ws + L.partial);
  }
  const int nblk = (npix + 255) / 256;
  combine_kernel<T><<<dim3(nblk, n_img), 256, 0, st>>>(cc, chr, lowc, s, out, partial);
  if ((rc = check_launch("combine_kernel"))) return rc;
  if (data) {
    logl_kernel<<<n_img, 64, 0, st>>>(partial, nblk, logl);
    return check_launch("logl_kernel");
  }
  return LENSIM_OK;
}

// Zero-padded copy of [n_img][ny][nx] into [n_img][ny + 2 cy][nx + 2 cx].
template <typename T>
__global__ void pad_kernel(const T *__restrict__ in, int nx, int ny, int cx, int cy, T *__restrict__ out) {
  const int nxe = nx + 2 * cx, nye = ny + 2 * cy;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nxe * nye) return;
  const int img = blockIdx.y;
  const int Y = idx / nxe, X = idx - Y * nxe;
  const int y = Y - cy, x = X - cx;
  const T v = (y >= 0 && y < ny && x >= 0 && x < nx) ? AT(in, size_t(img) * nx * ny + size_t(y) * nx + x) : T(0);
  AT(out, size_t(img) * nxe * nye + idx) = v;
}

// Adjoint of edge clamping: pixel (y, x) receives every extended pixel
// (Y, X) with clamp(Y - cy) = y and clamp(X - cx) = x.
template <typename T>
__global__ void fold_kernel(const T *__restrict__ ext, int nx, int ny, int cx, int cy, T *__restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nx * ny) return;
  const int img = blockIdx.y;
  const int y = idx / nx, x = idx - y * nx;
  const int nxe = nx + 2 * cx, nye = ny + 2 * cy;
  const int y0 = y == 0 ? 0 : y + cy, y1 = y == ny - 1 ? nye - 1 : y + cy;
  const int x0 = x == 0 ? 0 : x + cx, x1 = x == nx - 1 ? nxe - 1 : x + cx;
  const T *e = ext + size_t(img) * nxe * nye;
  double s = 0.0;
  for (int Y = y0; Y <= y1; ++Y)
    for (int X = x0; X <= x1; ++X) s += double(AT(e, size_t(Y) * nxe + X));
  AT(out, size_t(img) * nx * ny + idx) = T(s);
}

// Adjoint of the 'same' convolution of [n_img][ny][nx] by a kernel prepared
// UNflipped in kf (a correlation): zero padding directly; edge padding via the
// extended domain.  The forward pass with edge padding is a valid correlation
// of the edge-extended image E[Y][X] = in[clamp(Y - cy)][clamp(X - cx)], so its
// adjoint is the zero-padded 'same' correlation over the extended domain,
// folded back onto the pixels each extended one copies.
template <typename T>
int adjoint_conv(int nx, int ny, int kx, int ky, int edge, const T *gin, const double *kf, T *gout, T *gpad, T *gext,
                 int n_img, hipStream_t st) {
  if (!edge) return launch_conv_plain<T>(nx, ny, kx, ky, 0, gin, kf, gout, n_img, st);
  const int cx = (kx - 1) / 2, cy = (ky - 1) / 2;
  const int nxe = nx + 2 * cx, nye = ny + 2 * cy;
  const int ne = nxe * nye;
  pad_kernel<T><<<dim3((ne + 255) / 256, n_img), 256, 0, st>>>(gin, nx, ny, cx, cy, gpad);
  int rc = check_launch("pad_kernel");
  if (rc) return rc;
  if ((rc = launch_conv_plain<T>(nxe, nye, kx, ky, 0, gpad, kf, gext, n_img, st))) return rc;
  fold_kernel<T><<<dim3((nx * ny + 255) / 256, n_img), 256, 0, st>>>(gext, nx, ny, cx, cy, gout);
  return check_launch("fold_kernel");
}

// log-likelihood and its gradient: the forward pipeline writes dlogL/dmodel,
// the adjoint of the PSF (a correlation with the kernel) gives dlogL/d(ray
// image) -- through the s x s mean's adjoint first when the convolution is
// supersampled -- and grad_kernel accumulates d(ray flux)/d(kwargs) per
// profile over the grid the convolution saw.
template <typename T>
int run_grad(const ModelArgs &m, const GridArgs &g, const lensim_psf *psf, int kx, int ky, int edge,
             const SSPlan &ss, const lensim_data *data, const double *params, int n_img, double *logl, double *grad,
             unsigned char *ws, const Layout &L, hipStream_t st) {
  T *gmodel = reinterpret_cast<T *>(ws + L.gmodel);
  const unsigned flags = LENSIM_IMAGE_SOURCE | LENSIM_IMAGE_LENS_LIGHT;
  int rc = ss.on ? run_image_ss<T>(m, g, psf, kx, ky, edge, ss, data, params, n_img, flags, nullptr, logl, ws, L, st,
                                   gmodel)
                 : run_image<T>(m, g, psf, kx, ky, edge, data, params, n_img, flags, nullptr, logl, ws, L, st, gmodel);
  if (rc) return rc;
  const bool identity = !psf || psf->type == LENSIM_PSF_NONE;
  const int kyp = kyp_of(ky);
  double *kf = reinterpret_cast<double *>(ws + L.psf);
  psf_prep_kernel<<<(kyp * kx + 255) / 256, 256, 0, st>>>(identity ? nullptr : psf->d_kernel, kx, ky, kyp, kf, 0);
  if ((rc = check_launch("psf_prep_kernel"))) return rc;
  T *glow = reinterpret_cast<T *>(ws + L.glow);
  T *gpad = reinterpret_cast<T *>(ws + L.gpad);
  T *gext = reinterpret_cast<T *>(ws + L.gext);
  const GridArgs gg = ss.on ? ray_grid(g) : g;  // the grid whose pixels the rays fill
  if (!ss.on) {
    if ((rc = adjoint_conv<T>(g.nx, g.ny, kx, ky, edge, gmodel, kf, glow, gpad, gext, n_img, st))) return rc;
  } else {
    T *gchr = reinterpret_cast<T *>(ws + L.gchr);
    const int nh = gg.nx * gg.ny;
    unpool_kernel<T><<<dim3((nh + 255) / 256, n_img), 256, 0, st>>>(gmodel, g.nx, g.ny, ss.s, 0, gchr);
    if ((rc = check_launch("unpool_kernel"))) return rc;
    if ((rc = adjoint_conv<T>(gg.nx, gg.ny, kx, ky, edge, gchr, kf, glow, gpad, gext, n_img, st))) return rc;
    if (ss.kxl > 0) {  // low-resolution part: adjoint of its convolution, then of the mean
      const int kypl = kyp_of(ss.kyl);
      double *kfl = reinterpret_cast<double *>(ws + L.psf_low);
      psf_prep_kernel<<<(kypl * ss.kxl + 255) / 256, 256, 0, st>>>(psf->d_kernel_low, ss.kxl, ss.kyl, kypl, kfl, 0);
      if ((rc = check_launch("psf_prep_kernel (low)"))) return rc;
      T *glu = reinterpret_cast<T *>(ws + L.glu);
      if ((rc = launch_conv_plain<T>(g.nx, g.ny, ss.kxl, ss.kyl, 0, gmodel, kfl, glu, n_img, st))) return rc;
      unpool_kernel<T><<<dim3((nh + 255) / 256, n_img), 256, 0, st>>>(glu, g.nx, g.ny, ss.s, 1, glow);
      if ((rc = check_launch("unpool_kernel"))) return rc;
    }
  }
  // per-profile gradients
  GradArgs ga;
  ga.params = params;
  ga.coefd = reinterpret_cast<const double *>(ws + L.coefd);
  ga.coeff = reinterpret_cast<const float *>(ws + L.coeff);
  ga.dcoefd = reinterpret_cast<const double *>(ws + L.dcoefd);
  ga.dcoeff = reinterpret_cast<const float *>(ws + L.dcoeff);
  ga.glow = glow;
  ga.gpart = reinterpret_cast<double *>(ws + L.gpart);
  ga.gbeta = L.gbeta ? static_cast<void *>(ws + L.gbeta) : nullptr;
  const int npix = gg.nx * gg.ny, nblk = (npix + 256 * kGradPix - 1) / (256 * kGradPix);
  const double *recd = reinterpret_cast<const double *>(ws + L.recd);
  const float *recf = reinterpret_cast<const float *>(ws + L.recf);
  // multi-plane: every deflector in one launch, which also stores beta at every
  // ray (ga.gbeta) for the source launches
  if (m.mp_n && (rc = launch_mp_grad<T>(m, gg, ga, recd, recf, reinterpret_cast<double *>(ws + L.gpart_mp), nblk,
                                        n_img, st)))
    return rc;
  // the sources first: the first source's launch stores dS/dbeta at every ray
  // (ga.gbeta), which the inline lens launches after it read
  for (int q = 0; q < m.n_prof; ++q) {
    const int k = q < m.n_source ? m.n_defl + q : (q < m.n_source + m.n_defl ? q - m.n_source : q);
    if (m.mp_n && k < m.n_defl) continue;  // multi-plane: mp_grad_kernel above
    if (k == m.n_lens && m.n_bulk > 0) {  // the bulk list in one launch
      rc = m.has_halo ? launch_bulk_grad<T, true>(m, gg, ga, recd, recf, nblk, n_img, st)
                      : launch_bulk_grad<T, false>(m, gg, ga, recd, recf, nblk, n_img, st);
      if (rc) return rc;
      q += m.n_bulk - 1;
      continue;
    }
    rc = m.has_halo ? launch_grad_profile<T, true>(m, gg, ga, recd, recf, nblk, n_img, k, st)
                    : launch_grad_profile<T, false>(m, gg, ga, recd, recf, nblk, n_img, k, st);
    if (rc) return rc;
  }
  const int n = n_img * m.n_prof * LENSIM_NPARAM;
  grad_reduce_kernel<<<(n + 255) / 256, 256, 0, st>>>(ga.gpart, nblk, n, grad);
  return check_launch("grad_reduce_kernel");
}

// ------------------------------------------------- linear amplitude solve --
// ImageLinearFit.image_linear_solve (ImSim/image_linear_solve.py:98-181): the
// response of every linear light profile (its image at amp = 1,
// linear_response_matrix :362-455), weighted least squares against the data
// (de_lens.get_param_WLS, ImSim/de_lens.py:10-36) and the likelihood of the
// solved model (:229-340).
constexpr int kMaxLin = 8;  // linear amplitudes per model
constexpr int kLinAcc = kMaxLin * (kMaxLin + 1) / 2 + kMaxLin;  // M (upper triangle) and R

// params with the amplitude (slot 0) of every light profile set to 1
__global__ void amp_one_kernel(const double *__restrict__ params, int n_img, int n_prof, int k_light0,
                               double *__restrict__ out) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n_img * n_prof * LENSIM_NPARAM) return;
  const int q = idx % LENSIM_NPARAM, k = (idx / LENSIM_NPARAM) % n_prof;
  AT(out, idx) = (k >= k_light0 && q == 0) ? 1.0 : AT(params, idx);
}

// Weights mask / C_D(data) (ImageModel.error_response, image_model.py:577-596:
// Data.C_D, the variance estimated from the data) and per-block partial sums
// of M = A W A^T (upper triangle) and R = A W d.  A is [n_lin][n_img][npix].
template <typename T>
__global__ __launch_bounds__(256) void normal_kernel(ConvArgs c, const T *__restrict__ A, int n_lin, int n_img,
                                                     double *__restrict__ part) {
  const int img = blockIdx.y;
  const int npix = c.nx * c.ny;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double acc[kLinAcc];
#pragma unroll
  for (int q = 0; q < kLinAcc; ++q) acc[q] = 0.0;
  if (p < npix) {
    const size_t pd = p + size_t(img) * size_t(c.dstride);
    const double d = AT(c.data, pd);
    double var;
    if (c.noise) {
      var = AT(c.noise, pd) * AT(c.noise, pd);
    } else {
      const double t = c.expo ? AT(c.expo, p) : c.exposure_time;
      var = (d >= 0.0 ? d : 0.0) / t + c.bkg2;
    }
    const double w = (c.mask ? AT(c.mask, p) : 1.0) / var;
    double a[kMaxLin];
#pragma unroll
    for (int l = 0; l < kMaxLin; ++l) a[l] = l < n_lin ? double(AT(A, (size_t(l) * n_img + img) * npix + p)) : 0.0;
    if (w != 0.0) {
      int q = 0;
#pragma unroll
      for (int i = 0; i < kMaxLin; ++i) {
#pragma unroll
        for (int j = i; j < kMaxLin; ++j) acc[q++] = a[i] * w * a[j];
      }
#pragma unroll
      for (int i = 0; i < kMaxLin; ++i) acc[q++] = a[i] * w * d;
    }
  }
  __shared__ double red[4][kLinAcc];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < kLinAcc; ++q) {
    double v = acc[q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) red[wave][q] = v;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < kLinAcc; q += blockDim.x)
    AT(part, (size_t(img) * gridDim.x + blockIdx.x) * kLinAcc + q) = red[0][q] + red[1][q] + red[2][q] + red[3][q];
}

// Eigenvalues of a symmetric n x n matrix (cyclic Jacobi rotations).
__device__ void sym_eigenvalues(double (*a)[kMaxLin], int n, double *ev) {
  for (int sweep = 0; sweep < 30; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = i + 1; j < n; ++j) off += a[i][j] * a[i][j];
    if (off == 0.0) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (a[p][q] == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * a[p][q]);
        const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // columns p, q
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // rows p, q
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
      }
  }
  for (int i = 0; i < n; ++i) ev[i] = a[i][i];
}

// LU factorisation with partial pivoting in place; returns the permutation
// parity (+1 / -1), 0 if a pivot is exactly zero.
__device__ int lu_factor(double (*a)[kMaxLin], int n, int *piv) {
  int parity = 1;
  for (int k = 0; k < n; ++k) {
    int pr = k;
    for (int i = k + 1; i < n; ++i)
      if (fabs(a[i][k]) > fabs(a[pr][k])) pr = i;
    piv[k] = pr;
    if (pr != k) {
      parity = -parity;
      for (int j = 0; j < n; ++j) {
        const double t = a[k][j];
        a[k][j] = a[pr][j];
        a[pr][j] = t;
      }
    }
    if (a[k][k] == 0.0) return 0;
    for (int i = k + 1; i < n; ++i) {
      a[i][k] /= a[k][k];
      for (int j = k + 1; j < n; ++j) a[i][j] -= a[i][k] * a[k][j];
    }
  }
  return parity;
}

__device__ void lu_solve(const double (*lu)[kMaxLin], const int *piv, int n, double *b) {
  for (int k = 0; k < n; ++k) {
    const double t = b[k];
    b[k] = b[piv[k]];
    b[piv[k]] = t;
  }
  for (int i = 1; i < n; ++i)
    for (int j = 0; j < i; ++j) b[i] -= lu[i][j] * b[j];
  for (int i = n - 1; i >= 0; --i) {
    for (int j = i + 1; j < n; ++j) b[i] -= lu[i][j] * b[j];
    b[i] /= lu[i][i];
  }
}

// One thread per image: sums the partials, checks cond(M) < 5 / eps, solves
// M B = R (B = 0 when unstable or not finite), optionally inverts M
// (nan_to_num'ed, zero when unstable) and the marginalisation constant
// (de_lens.marginalisation_const / marginalization_new, de_lens.py:40-77).
__global__ void lin_solve_kernel(const double *__restrict__ part, int nblk, int n_lin, int n_img, int source_marg,
                                 double d_prior, double *__restrict__ amp, double *__restrict__ cov,
                                 double *__restrict__ marg) {
  const int img = blockIdx.x * blockDim.x + threadIdx.x;
  if (img >= n_img) return;
  double acc[kLinAcc];
  for (int q = 0; q < kLinAcc; ++q) acc[q] = 0.0;
  for (int b = 0; b < nblk; ++b)
    for (int q = 0; q < kLinAcc; ++q) acc[q] += AT(part, (size_t(img) * nblk + b) * kLinAcc + q);
  double M[kMaxLin][kMaxLin], W[kMaxLin][kMaxLin], R[kMaxLin];
  int q = 0;
  for (int i = 0; i < kMaxLin; ++i)
    for (int j = i; j < kMaxLin; ++j) {
      M[i][j] = M[j][i] = acc[q];
      ++q;
    }
  for (int i = 0; i < kMaxLin; ++i) R[i] = acc[q++];
  const int n = n_lin;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) W[i][j] = M[i][j];
  double ev[kMaxLin];
  sym_eigenvalues(W, n, ev);
  double emax = 0.0, emin = INFINITY;
  for (int i = 0; i < n; ++i) {
    emax = fmax(emax, fabs(ev[i]));
    emin = fmin(emin, fabs(ev[i]));
  }
  const double eps = 2.220446049250313e-16;
  const bool stable = emin > 0.0 && emax / emin < 5.0 / eps;
  int piv[kMaxLin];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) W[i][j] = M[i][j];
  const int parity = lu_factor(W, n, piv);
  double B[kMaxLin];
  for (int i = 0; i < n; ++i) B[i] = R[i];
  bool finite = stable && parity != 0;
  if (finite) {
    lu_solve(W, piv, n, B);
    for (int i = 0; i < n; ++i) finite = finite && isfinite(B[i]);
  }
  for (int i = 0; i < n; ++i) AT(amp, size_t(img) * n + i) = finite ? B[i] : 0.0;
  if (cov) {
    for (int j = 0; j < n; ++j) {
      double e[kMaxLin];
      for (int i = 0; i < n; ++i) e[i] = i == j ? 1.0 : 0.0;
      if (stable && parity != 0) lu_solve(W, piv, n, e);
      for (int i = 0; i < n; ++i) {
        const double v = (stable && parity != 0 && isfinite(e[i])) ? e[i] : 0.0;
        AT(cov, (size_t(img) * n + i) * n + j) = v;
      }
    }
  }
  double mc = 0.0;
  if (source_marg) {
    if (!(stable && parity != 0)) {
      mc = -1e15;  // slogdet / eig of a zero covariance
    } else if (!(d_prior > 0.0)) {
      // 1/2 log det(M^-1) = -1/2 sum log|U_ii| with the sign of det(M)
      double logdet = 0.0;
      int sign = parity;
      for (int i = 0; i < n; ++i) {
        logdet += log(fabs(W[i][i]));
        if (W[i][i] < 0.0) sign = -sign;
      }
      mc = sign * (-logdet) / 2.0;
    } else {
      // eigenvalues of M^-1 = 1 / eigenvalues of M, clipped at d_prior^2
      double ld = 0.0, sgn = 1.0;
      for (int i = 0; i < n; ++i) {
        const double v = 1.0 / ev[i];
        sgn *= v > 0.0 ? 1.0 : (v < 0.0 ? -1.0 : 0.0);
        ld += log(fmin(fabs(v), d_prior * d_prior));
      }
      const double log_det = ld * sgn;
      mc = isnan(log_det) ? -1e15 : log_det / 2.0 + n / 2.0 * log(M_PI / 2.0) - n * log(d_prior);
    }
  }
  AT(marg, img) = mc;
}

// Solved model (zero outside the likelihood mask, array_masked2image) and its chi^2.
template <typename T>
__global__ __launch_bounds__(256) void lin_model_kernel(ConvArgs c, const T *__restrict__ A, const double *__restrict__ amp,
                                                        int n_lin, int n_img, T *__restrict__ out,
                                                        double *__restrict__ partial) {
  const int img = blockIdx.y;
  const int npix = c.nx * c.ny;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double chi = 0.0;
  if (p < npix) {
    double m = 0.0;
    if (!c.mask || AT(c.mask, p) != 0.0)
      for (int l = 0; l < n_lin; ++l)
        m += AT(amp, size_t(img) * n_lin + l) * double(AT(A, (size_t(l) * n_img + img) * npix + p));
    const T mt = T(m);
    if (out) AT(out, size_t(img) * npix + p) = mt;
    chi = chi2_pixel<T>(c, p, mt, img);
  }
  __shared__ double red[4];
  for (int off = 32; off > 0; off >>= 1) chi += __shfl_xor(chi, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = chi;
  __syncthreads();
  if (threadIdx.x == 0) AT(partial, size_t(img) * gridDim.x + blockIdx.x) = red[0] + red[1] + red[2] + red[3];
}

__global__ void add_kernel(double *__restrict__ x, const double *__restrict__ y, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) AT(x, i) += AT(y, i);
}

// Workspace of the linear solve: the image layout of one column, then the
// columns A, the amp = 1 parameters, partials, amplitudes and constants.
struct LinLayout {
  Layout img;
  size_t A, params1, part, chi, amp, marg, total;
};

LinLayout lin_layout(const ModelArgs &m, int nx, int ny, int kx, int ky, int n_img, int precision, const SSPlan &ss,
                     int n_lin) {
  LinLayout L;
  L.img = layout(m, nx, ny, kx, ky, n_img, precision, ss);
  const size_t tsz = precision == LENSIM_MIXED ? 4 : 8;
  const size_t npix = size_t(nx) * ny, nblk = (npix + 255) / 256;
  size_t off = L.img.total;
  L.A = off;
  off = align256(off + size_t(n_lin) * n_img * npix * tsz);
  L.params1 = off;
  off = align256(off + size_t(n_img) * m.n_prof * LENSIM_NPARAM * 8);
  L.part = off;
  off = align256(off + size_t(n_img) * nblk * kLinAcc * 8);
  L.chi = off;
  off = align256(off + size_t(n_img) * nblk * 8);
  L.amp = off;
  off = align256(off + size_t(n_img) * kMaxLin * 8);
  L.marg = off;
  off = align256(off + size_t(n_img) * 8);
  L.total = off;
  return L;
}

template <typename T>
int run_linear(const ModelArgs &m, const GridArgs &g, const lensim_psf *psf, int kx, int ky, int edge,
               const SSPlan &ss, const lensim_data *data, const double *params, int n_img, int source_marg,
               double d_prior, double *amp_out, double *logl, T *model_out, double *cov, T *resp_out,
               unsigned char *ws, const LinLayout &LL, hipStream_t st) {
  const int n_lin = m.n_source + m.n_light;
  const int npix = g.nx * g.ny, nblk = (npix + 255) / 256;
  double *params1 = reinterpret_cast<double *>(ws + LL.params1);
  const int ntot = n_img * m.n_prof * LENSIM_NPARAM;
  amp_one_kernel<<<(ntot + 255) / 256, 256, 0, st>>>(params, n_img, m.n_prof, m.n_defl, params1);
  int rc = check_launch("amp_one_kernel");
  if (rc) return rc;
  T *A = resp_out ? resp_out : reinterpret_cast<T *>(ws + LL.A);
  // one column per linear profile: the model with every other light profile off
  for (int l = 0; l < n_lin; ++l) {
    ModelArgs mc = m;
    for (int k = m.n_lens; k < m.n_lens + n_lin; ++k)
      if (k - m.n_lens != l) mc.types[k] = LENSIM_NULL;
    const unsigned flags = l < m.n_source ? LENSIM_IMAGE_SOURCE : LENSIM_IMAGE_LENS_LIGHT;
    T *col = A + size_t(l) * n_img * npix;
    rc = ss.on ? run_image_ss<T>(mc, g, psf, kx, ky, edge, ss, nullptr, params1, n_img, flags, col, nullptr, ws,
                                 LL.img, st)
               : run_image<T>(mc, g, psf, kx, ky, edge, nullptr, params1, n_img, flags, col, nullptr, ws, LL.img, st);
    if (rc) return rc;
  }
  ConvArgs c{};
  c.nx = g.nx;
  c.ny = g.ny;
  c.data = data->d_data;
  c.noise = data->d_noise_map;
  c.expo = data->d_exposure_map;
  c.mask = data->d_mask;
  c.exposure_time = data->exposure_time;
  c.bkg2 = data->background_rms * data->background_rms;
  c.inv_t = 1.0 / data->exposure_time;
  c.dstride = data->image_stride;
  double *part = reinterpret_cast<double *>(ws + LL.part);
  normal_kernel<T><<<dim3(nblk, n_img), 256, 0, st>>>(c, A, n_lin, n_img, part);
  if ((rc = check_launch("normal_kernel"))) return rc;
  double *amp = reinterpret_cast<double *>(ws + LL.amp);
  double *marg = reinterpret_cast<double *>(ws + LL.marg);
  lin_solve_kernel<<<(n_img + 63) / 64, 64, 0, st>>>(part, nblk, n_lin, n_img, source_marg, d_prior, amp, cov, marg);
  if ((rc = check_launch("lin_solve_kernel"))) return rc;
  double *chi = reinterpret_cast<double *>(ws + LL.chi);
  lin_model_kernel<T><<<dim3(nblk, n_img), 256, 0, st>>>(c, A, amp, n_lin, n_img, model_out, chi);
  if ((rc = check_launch("lin_model_kernel"))) return rc;
  logl_kernel<<<n_img, 64, 0, st>>>(chi, nblk, logl);
  if ((rc = check_launch("logl_kernel"))) return rc;
  add_kernel<<<(n_img + 255) / 256, 256, 0, st>>>(logl, marg, n_img);
  if ((rc = check_launch("add_kernel"))) return rc;
  if (amp_out) {
    const hipError_t e = hipMemcpyAsync(amp_out, amp, size_t(n_img) * n_lin * 8, hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(amp)");
  }
  return LENSIM_OK;
}

// ------------------------------------------- allocation-free entry point --
// lensim_loglike_noalloc keeps its scratch memory in this file-scope device
// array (one copy per GPU, held until the process ends) instead of a
// workspace argument.  In host code `scratch` names a host-side stand-in, not
// the device array: its device address on the current GPU comes from
// hipGetSymbolAddress (scratch_on_device) and only that address reaches
// kernels.  No kernel of the log-likelihood pipeline reads scratch memory it
// has not written earlier in the same call, so nothing is cleared between
// calls.  Calls that overlap on one GPU share the array and race.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current GPU, looked up once per GPU.
int scratch_on_device(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(LENSIM_EINVAL, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(LENSIM_EDEVICE, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LENSIM_OK;
}

int check_ws(const Layout &L, void *ws, size_t ws_bytes) {
  if (L.total > ws_bytes)
    return fail(LENSIM_EWORKSPACE, "workspace of %zu B is smaller than the %zu B needed", ws_bytes, L.total);
  if (L.total > 0 && !ws) return fail(LENSIM_EINVAL, "workspace pointer is NULL");
  if (reinterpret_cast<uintptr_t>(ws) & 255) return fail(LENSIM_EINVAL, "workspace must be 256-byte aligned");
  return LENSIM_OK;
}

// The caller's buffers that a call reads: parameters, the bulk type list, the
// multi-plane arrays, the evaluation mask, the PSF kernels and the data, each
// with the extent the header gives it.
void add_inputs(CallRanges &cr, const lensim_model *model, const ModelArgs &m, const double *params, int n_img,
                const lensim_grid *grid, const lensim_psf *psf, const lensim_data *data, int nx, int ny) {
  cr.add(params, size_t(n_img) * m.n_prof * LENSIM_NPARAM * sizeof(double));
  if (m.bulk_types) cr.add(m.bulk_types, size_t(m.n_bulk) * sizeof(int32_t));
  if (const lensim_multiplane *mp = model->multiplane) {
    const size_t n = mp->stride ? size_t(n_img - 1) * size_t(mp->stride) + mp->n_planes : size_t(mp->n_planes);
    cr.add(mp->d_order, size_t(mp->n_planes - 1) * sizeof(int32_t));
    cr.add(mp->d_T_ij, n * sizeof(double));
    cr.add(mp->d_T_z, n * sizeof(double));
    cr.add(mp->d_r2p, n * sizeof(double));
  }
  if (grid) cr.add(grid->d_eval_mask, size_t(grid->nx) * grid->ny);
  if (psf) {
    cr.add(psf->d_kernel, size_t(psf->kx) * psf->ky * sizeof(double));
    cr.add(psf->d_kernel_low, size_t(psf->kx_low) * psf->ky_low * sizeof(double));
  }
  if (data) {
    const size_t npix = size_t(nx) * ny;
    const size_t nd = data->image_stride ? size_t(n_img - 1) * size_t(data->image_stride) + npix : npix;
    cr.add(data->d_data, nd * sizeof(double));
    cr.add(data->d_noise_map, nd * sizeof(double));
    cr.add(data->d_exposure_map, npix * sizeof(double));
    cr.add(data->d_mask, npix * sizeof(double));
  }
}

#ifdef LENSIM_BOUNDS_CHECK
// Stores bytes 0 .. n - 1 of buf and then byte n (one past the registered range).
__global__ void bounds_selftest_kernel(unsigned char *buf, size_t n) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i <= n) AT(buf, i) = static_cast<unsigned char>(i & 255);
}
#endif

}  // namespace

// ------------------------------------------------------------- C ABI -----

extern "C" {

int lensim_abi_version(void) { return LENSIM_ABI_VERSION; }

int lensim_timing_enable(int on) {
  g_timing = on != 0;
  return LENSIM_OK;
}

int lensim_timing_read(double *render_ms, double *conv_ms, int64_t *calls) {
  double r = 0.0, c = 0.0;
  int64_t n = 0;
  for (int i = 0; i < g_ncalls; ++i) {
    TimedCall &tc = g_calls[i];
    if (!tc.ev[0] || !tc.ev[1] || !tc.ev[2]) continue;
    hipError_t e = hipEventSynchronize(tc.ev[2]);
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, tc.ev[0], tc.ev[1]) != hipSuccess ||
        hipEventElapsedTime(&b, tc.ev[1], tc.ev[2]) != hipSuccess)
      return fail(LENSIM_EDEVICE, "hipEventElapsedTime failed");
    r += a;
    c += b;
    ++n;
  }
  g_ncalls = 0;  // events are kept and reused by later calls
  if (render_ms) *render_ms = r;
  if (conv_ms) *conv_ms = c;
  if (calls) *calls = n;
  return LENSIM_OK;
}

const char *lensim_last_error(void) { return g_err; }

int lensim_bounds_checked(void) {
#ifdef LENSIM_BOUNDS_CHECK
  return 1;
#else
  return 0;
#endif
}

int lensim_bounds_selftest(void *d_buf, size_t bytes, void *stream) {
#ifdef LENSIM_BOUNDS_CHECK
  if (!d_buf) return fail(LENSIM_EINVAL, "NULL buffer");
  if (bytes == 0 || bytes > (size_t(1) << 20)) return fail(LENSIM_EINVAL, "bytes must be in [1, 2^20]");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  CallRanges cr;
  cr.add(d_buf, bytes);
  int rc = cr.begin(st);
  if (rc) return rc;
  bounds_selftest_kernel<<<unsigned((bytes + 1 + 255) / 256), 256, 0, st>>>(static_cast<unsigned char *>(d_buf),
                                                                             bytes);
  return cr.end(st, check_launch("bounds_selftest_kernel"));
#else
  (void)d_buf;
  (void)bytes;
  (void)stream;
  return fail(LENSIM_EUNSUPPORTED, "lensim_bounds_selftest needs the bounds-checked build (liblensim_checked.so)");
#endif
}

size_t lensim_workspace_size(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                             int32_t n_img, int32_t precision) {
  ModelArgs m;
  if (model_args(model, m)) return 0;
  int nx = 1, ny = 1;
  SSPlan ss;
  if (grid) {
    nx = grid->nx;
    ny = grid->ny;
    GridArgs g;
    if (grid_args(grid, g) || psf_ss(psf, g, 0, ss)) return 0;
  }
  int kx = 1, ky = 1, edge = 0;
  if (psf_dims(psf, 0, kx, ky, edge)) return 0;
  return layout(m, nx, ny, kx, ky, n_img, precision, ss).total;
}

int lensim_raytrace(const lensim_model *model, const double *d_params, int32_t n_img, const double *d_x,
                    const double *d_y, int64_t n_ray, int32_t precision, double *d_beta_x, double *d_beta_y,
                    void *ws, size_t ws_bytes, void *stream) {
  ModelArgs m;
  int rc = model_args(model, m);
  if (rc) return rc;
  if (n_img < 0 || n_ray < 0) return fail(LENSIM_EINVAL, "negative sizes");
  if (n_img == 0 || n_ray == 0) return LENSIM_OK;
  if (n_ray > (int64_t(1) << 36)) return fail(LENSIM_EINVAL, "n_ray too large");
  if (n_img > 65535) return fail(LENSIM_EINVAL, "n_img %d exceeds 65535 per call", n_img);
  if (!d_x || !d_y || !d_beta_x || !d_beta_y || (m.n_prof && !d_params))
    return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  const Layout L = layout(m, 1, 1, 1, 1, n_img, precision);
  if ((rc = check_ws(L, ws, ws_bytes))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char *w = static_cast<unsigned char *>(ws);
  CallRanges cr;
  add_inputs(cr, model, m, d_params, n_img, nullptr, nullptr, nullptr, 0, 0);
  cr.add(d_x, size_t(n_ray) * sizeof(double));
  cr.add(d_y, size_t(n_ray) * sizeof(double));
  cr.add(d_beta_x, size_t(n_img) * size_t(n_ray) * sizeof(double));
  cr.add(d_beta_y, size_t(n_img) * size_t(n_ray) * sizeof(double));
  cr.add(ws, L.total);
  if ((rc = cr.begin(st))) return rc;
  if ((rc = launch_prep(m, d_params, n_img, w, L, st))) return cr.end(st, rc);
  dim3 grid((unsigned)((n_ray + 255) / 256), n_img);
  const double *recd = reinterpret_cast<const double *>(w + L.recd);
  const float *recf = reinterpret_cast<const float *>(w + L.recf);
  const double *coefd = reinterpret_cast<const double *>(w + L.coefd);
  const float *coeff = reinterpret_cast<const float *>(w + L.coeff);
  const dim3 grid_mp((unsigned)((n_ray * kMpG + 255) / 256), n_img);
#define LENSIM_TRACE(T, H, D)                                                                                      \
  do {                                                                                                             \
    if (m.mp_n && sizeof(T) == sizeof(float))                                                                      \
      raytrace_mp_kernel<T, H, D><<<grid_mp, 256, 0, st>>>(m, recd, recf, coefd, coeff, d_x, d_y, n_ray, d_beta_x, \
                                                           d_beta_y);                                             \
    else                                                                                                           \
      raytrace_kernel<T, H, D><<<grid, 256, 0, st>>>(m, recd, recf, coefd, coeff, d_x, d_y, n_ray, d_beta_x,       \
                                                     d_beta_y);                                                   \
  } while (0)
  if (precision == LENSIM_FP64) {
    if (m.has_halo) LENSIM_TRACE(double, true, false);
    else LENSIM_TRACE(double, false, false);
  } else if (m.has_nie_nfw) {  // SIE / NIE and NFW deflect in double (see nie_defl)
    if (m.has_halo) LENSIM_TRACE(float, true, true);
    else LENSIM_TRACE(float, false, true);
  } else {
    if (m.has_halo) LENSIM_TRACE(float, true, false);
    else LENSIM_TRACE(float, false, false);
  }
#undef LENSIM_TRACE
  return cr.end(st, check_launch("raytrace_kernel"));
}

int lensim_surface_brightness(const lensim_model *model, const double *d_params, int32_t n_img, int32_t which,
                              const double *d_x, const double *d_y, int64_t n_ray, int32_t precision,
                              double *d_flux, void *ws, size_t ws_bytes, void *stream) {
  ModelArgs m;
  int rc = model_args(model, m);
  if (rc) return rc;
  if (n_img < 0 || n_ray < 0) return fail(LENSIM_EINVAL, "negative sizes");
  if (which != 0 && which != 1) return fail(LENSIM_EINVAL, "which must be 0 (source) or 1 (lens light)");
  if (n_img == 0 || n_ray == 0) return LENSIM_OK;
  if (n_ray > (int64_t(1) << 36)) return fail(LENSIM_EINVAL, "n_ray too large");
  if (n_img > 65535) return fail(LENSIM_EINVAL, "n_img %d exceeds 65535 per call", n_img);
  if (!d_x || !d_y || !d_flux || (m.n_prof && !d_params)) return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  const Layout L = layout(m, 1, 1, 1, 1, n_img, precision);
  if ((rc = check_ws(L, ws, ws_bytes))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char *w = static_cast<unsigned char *>(ws);
  CallRanges cr;
  add_inputs(cr, model, m, d_params, n_img, nullptr, nullptr, nullptr, 0, 0);
  cr.add(d_x, size_t(n_ray) * sizeof(double));
  cr.add(d_y, size_t(n_ray) * sizeof(double));
  cr.add(d_flux, size_t(n_img) * size_t(n_ray) * sizeof(double));
  cr.add(ws, L.total);
  if ((rc = cr.begin(st))) return rc;
  if ((rc = launch_prep(m, d_params, n_img, w, L, st))) return cr.end(st, rc);
  const int k0 = which == 0 ? m.n_defl : m.n_defl + m.n_source;
  const int k1 = which == 0 ? m.n_defl + m.n_source : m.n_prof;
  dim3 grid((unsigned)((n_ray + 255) / 256), n_img);
  if (precision == LENSIM_FP64)
    light_kernel<double><<<grid, 256, 0, st>>>(m, reinterpret_cast<double *>(w + L.recd),
                                                reinterpret_cast<float *>(w + L.recf), k0, k1, d_x, d_y, n_ray,
                                                d_flux);
  else
    light_kernel<float><<<grid, 256, 0, st>>>(m, reinterpret_cast<double *>(w + L.recd),
                                               reinterpret_cast<float *>(w + L.recf), k0, k1, d_x, d_y, n_ray,
                                               d_flux);
  return cr.end(st, check_launch("light_kernel"));
}

int lensim_image(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf, const double *d_params,
                 int32_t n_img, int32_t precision, uint32_t flags, void *d_out, void *ws, size_t ws_bytes,
                 void *stream) {
  ModelArgs m;
  int rc = model_args(model, m);
  if (rc) return rc;
  GridArgs g;
  if ((rc = grid_args(grid, g))) return rc;
  int kx, ky, edge;
  if ((rc = psf_dims(psf, flags, kx, ky, edge))) return rc;
  if (n_img < 0) return fail(LENSIM_EINVAL, "negative n_img");
  if (n_img == 0) return LENSIM_OK;
  if (n_img > 65535) return fail(LENSIM_EINVAL, "n_img %d exceeds 65535 per call", n_img);
  if (!d_out || (m.n_prof && !d_params)) return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  // kernel dims used by the layout: the real PSF unless unconvolved
  int lkx = 1, lky = 1, le = 0;
  psf_dims(psf, 0, lkx, lky, le);
  SSPlan ss, ss_layout;
  if ((rc = psf_ss(psf, g, flags, ss))) return rc;
  if ((rc = psf_ss(psf, g, 0, ss_layout))) return rc;
  const Layout L = layout(m, g.nx, g.ny, lkx, lky, n_img, precision, ss_layout);
  if ((rc = check_ws(L, ws, ws_bytes))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char *w = static_cast<unsigned char *>(ws);
  const size_t tsz = precision == LENSIM_FP64 ? sizeof(double) : sizeof(float);
  CallRanges cr;
  add_inputs(cr, model, m, d_params, n_img, grid, psf, nullptr, g.nx, g.ny);
  cr.add(d_out, size_t(n_img) * g.nx * g.ny * tsz);
  cr.add(ws, L.total);
  if ((rc = cr.begin(st))) return rc;
  if (ss.on) {
    if (precision == LENSIM_FP64)
      rc = run_image_ss<double>(m, g, psf, kx, ky, edge, ss, nullptr, d_params, n_img, flags,
                                static_cast<double *>(d_out), nullptr, w, L, st);
    else
      rc = run_image_ss<float>(m, g, psf, kx, ky, edge, ss, nullptr, d_params, n_img, flags,
                               static_cast<float *>(d_out), nullptr, w, L, st);
  } else if (precision == LENSIM_FP64) {
    rc = run_image<double>(m, g, psf, kx, ky, edge, nullptr, d_params, n_img, flags, static_cast<double *>(d_out),
                           nullptr, w, L, st);
  } else {
    rc = run_image<float>(m, g, psf, kx, ky, edge, nullptr, d_params, n_img, flags, static_cast<float *>(d_out),
                          nullptr, w, L, st);
  }
  return cr.end(st, rc);
}

int lensim_loglike(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                   const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                   double *d_logl, void *d_model_out, void *ws, size_t ws_bytes, void *stream) {
  ModelArgs m;
  int rc = model_args(model, m);
  if (rc) return rc;
  GridArgs g;
  if ((rc = grid_args(grid, g))) return rc;
  int kx, ky, edge;
  if ((rc = psf_dims(psf, 0, kx, ky, edge))) return rc;
  if (!data || !data->d_data) return fail(LENSIM_EINVAL, "data image is NULL");
  if (!data->d_noise_map && !data->d_exposure_map && !(data->exposure_time > 0.0))
    return fail(LENSIM_EINVAL, "need a noise map, an exposure map or a positive exposure time");
  if (data->image_stride != 0 && data->image_stride < int64_t(g.nx) * g.ny)
    return fail(LENSIM_EINVAL, "image_stride %lld is smaller than an image", (long long)data->image_stride);
  if (n_img < 0) return fail(LENSIM_EINVAL, "negative n_img");
  if (n_img == 0) return LENSIM_OK;
  if (n_img > 65535) return fail(LENSIM_EINVAL, "n_img %d exceeds 65535 per call", n_img);
  if (!d_logl || (m.n_prof && !d_params)) return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  SSPlan ss;
  if ((rc = psf_ss(psf, g, 0, ss))) return rc;
  const Layout L = layout(m, g.nx, g.ny, kx, ky, n_img, precision, ss);
  if ((rc = check_ws(L, ws, ws_bytes))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char *w = static_cast<unsigned char *>(ws);
  const unsigned flags = LENSIM_IMAGE_SOURCE | LENSIM_IMAGE_LENS_LIGHT;
  const size_t tsz = precision == LENSIM_FP64 ? sizeof(double) : sizeof(float);
  CallRanges cr;
  add_inputs(cr, model, m, d_params, n_img, grid, psf, data, g.nx, g.ny);
  cr.add(d_logl, size_t(n_img) * sizeof(double));
  cr.add(d_model_out, size_t(n_img) * g.nx * g.ny * tsz);
  cr.add(ws, L.total);
  if ((rc = cr.begin(st))) return rc;
  if (ss.on) {
    if (precision == LENSIM_FP64)
      rc = run_image_ss<double>(m, g, psf, kx, ky, edge, ss, data, d_params, n_img, flags,
                                static_cast<double *>(d_model_out), d_logl, w, L, st);
    else
      rc = run_image_ss<float>(m, g, psf, kx, ky, edge, ss, data, d_params, n_img, flags,
                               static_cast<float *>(d_model_out), d_logl, w, L, st);
  } else if (precision == LENSIM_FP64) {
    rc = run_image<double>(m, g, psf, kx, ky, edge, data, d_params, n_img, flags, static_cast<double *>(d_model_out),
                           d_logl, w, L, st);
  } else {
    rc = run_image<float>(m, g, psf, kx, ky, edge, data, d_params, n_img, flags, static_cast<float *>(d_model_out),
                          d_logl, w, L, st);
  }
  return cr.end(st, rc);
}

size_t lensim_data_loglike_workspace_size(int32_t nx, int32_t ny, int32_t n_img) {
  if (nx <= 0 || ny <= 0 || n_img < 0) return 0;
  return align256(size_t(n_img) * ((size_t(nx) * ny + 255) / 256) * 8) + 256;
}

int lensim_data_loglike(const lensim_data *data, int32_t nx, int32_t ny, const void *d_model, int32_t n_img,
                        int32_t precision, double *d_logl, void *ws, size_t ws_bytes, void *stream) {
  if (nx <= 0 || ny <= 0) return fail(LENSIM_EINVAL, "image sizes must be positive (nx=%d ny=%d)", nx, ny);
  if (int64_t(nx) * ny > (int64_t(1) << 30)) return fail(LENSIM_EINVAL, "image too large");
  if (!data || !data->d_data) return fail(LENSIM_EINVAL, "data image is NULL");
  if (!data->d_noise_map && !data->d_exposure_map && !(data->exposure_time > 0.0))
    return fail(LENSIM_EINVAL, "need a noise map, an exposure map or a positive exposure time");
  if (data->image_stride != 0 && data->image_stride < int64_t(nx) * ny)
    return fail(LENSIM_EINVAL, "image_stride %lld is smaller than an image", (long long)data->image_stride);
  if (n_img < 0) return fail(LENSIM_EINVAL, "negative n_img");
  if (n_img == 0) return LENSIM_OK;
  if (n_img > 65535) return fail(LENSIM_EINVAL, "n_img %d exceeds 65535 per call", n_img);
  if (!d_model || !d_logl) return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  const size_t need = lensim_data_loglike_workspace_size(nx, ny, n_img);
  if (ws_bytes < need) return fail(LENSIM_EWORKSPACE, "workspace of %zu B is smaller than the %zu B needed", ws_bytes,
                                   need);
  if (!ws || (reinterpret_cast<uintptr_t>(ws) & 255)) return fail(LENSIM_EINVAL, "workspace must be 256-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  ConvArgs c{};
  c.nx = nx;
  c.ny = ny;
  c.data = data->d_data;
  c.noise = data->d_noise_map;
  c.expo = data->d_exposure_map;
  c.mask = data->d_mask;
  c.exposure_time = data->exposure_time;
  c.bkg2 = data->background_rms * data->background_rms;
  c.inv_t = 1.0 / data->exposure_time;
  c.dstride = data->image_stride;
  double *partial = static_cast<double *>(ws);
  const int nblk = (nx * ny + 255) / 256;
  CallRanges cr;
  const size_t npix = size_t(nx) * ny;
  const size_t nd = data->image_stride ? size_t(n_img - 1) * size_t(data->image_stride) + npix : npix;
  cr.add(data->d_data, nd * sizeof(double));
  cr.add(data->d_noise_map, nd * sizeof(double));
  cr.add(data->d_exposure_map, npix * sizeof(double));
  cr.add(data->d_mask, npix * sizeof(double));
  cr.add(d_model, size_t(n_img) * npix * (precision == LENSIM_FP64 ? sizeof(double) : sizeof(float)));
  cr.add(d_logl, size_t(n_img) * sizeof(double));
  cr.add(ws, need);
  int rc = cr.begin(st);
  if (rc) return rc;
  if (precision == LENSIM_FP64)
    data_chi2_kernel<double><<<dim3(nblk, n_img), 256, 0, st>>>(c, static_cast<const double *>(d_model), partial);
  else
    data_chi2_kernel<float><<<dim3(nblk, n_img), 256, 0, st>>>(c, static_cast<const float *>(d_model), partial);
  if ((rc = check_launch("data_chi2_kernel"))) return cr.end(st, rc);
  logl_kernel<<<n_img, 64, 0, st>>>(partial, nblk, d_logl);
  return cr.end(st, check_launch("logl_kernel"));
}

size_t lensim_linear_workspace_size(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                                    int32_t n_img, int32_t precision) {
  ModelArgs m;
  if (model_args(model, m) || !grid) return 0;
  GridArgs g;
  SSPlan ss;
  int kx = 1, ky = 1, edge = 0;
  if (grid_args(grid, g) || psf_ss(psf, g, 0, ss) || psf_dims(psf, 0, kx, ky, edge)) return 0;
  return lin_layout(m, g.nx, g.ny, kx, ky, n_img, precision, ss, m.n_source + m.n_light).total;
}

int lensim_linear_solve(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                        const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                        int32_t source_marg, double linear_prior, double *d_amp, double *d_logl, void *d_model_out,
                        double *d_cov, void *d_response_out, void *ws, size_t ws_bytes, void *stream) {
  ModelArgs m;
  int rc = model_args(model, m);
  if (rc) return rc;
  GridArgs g;
  if ((rc = grid_args(grid, g))) return rc;
  int kx, ky, edge;
  if ((rc = psf_dims(psf, 0, kx, ky, edge))) return rc;
  SSPlan ss;
  if ((rc = psf_ss(psf, g, 0, ss))) return rc;
  const int n_lin = m.n_source + m.n_light;
  if (n_lin == 0) return fail(LENSIM_EINVAL, "the linear solve needs at least one light profile");
  if (n_lin > kMaxLin) return fail(LENSIM_EUNSUPPORTED, "at most %d linear amplitudes (got %d)", kMaxLin, n_lin);
  if (!data || !data->d_data) return fail(LENSIM_EINVAL, "data image is NULL");
  if (!data->d_noise_map && !data->d_exposure_map && !(data->exposure_time > 0.0))
    return fail(LENSIM_EINVAL, "need a noise map, an exposure map or a positive exposure time");
  if (data->image_stride != 0 && data->image_stride < int64_t(g.nx) * g.ny)
    return fail(LENSIM_EINVAL, "image_stride %lld is smaller than an image", (long long)data->image_stride);
  if (n_img < 0) return fail(LENSIM_EINVAL, "negative n_img");
  if (n_img == 0) return LENSIM_OK;
  if (n_img > 65535) return fail(LENSIM_EINVAL, "n_img %d exceeds 65535 per call", n_img);
  if (!d_logl || !d_params) return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  const LinLayout LL = lin_layout(m, g.nx, g.ny, kx, ky, n_img, precision, ss, n_lin);
  if (LL.total > ws_bytes)
    return fail(LENSIM_EWORKSPACE, "workspace of %zu B is smaller than the %zu B needed", ws_bytes, LL.total);
  if (!ws || (reinterpret_cast<uintptr_t>(ws) & 255)) return fail(LENSIM_EINVAL, "workspace must be 256-byte aligned");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char *w = static_cast<unsigned char *>(ws);
  const size_t tsz = precision == LENSIM_FP64 ? sizeof(double) : sizeof(float);
  const size_t npix = size_t(g.nx) * g.ny;
  CallRanges cr;
  add_inputs(cr, model, m, d_params, n_img, grid, psf, data, g.nx, g.ny);
  cr.add(d_amp, size_t(n_img) * n_lin * sizeof(double));
  cr.add(d_logl, size_t(n_img) * sizeof(double));
  cr.add(d_model_out, size_t(n_img) * npix * tsz);
  cr.add(d_cov, size_t(n_img) * n_lin * n_lin * sizeof(double));
  cr.add(d_response_out, size_t(n_lin) * n_img * npix * tsz);
  cr.add(ws, LL.total);
  if ((rc = cr.begin(st))) return rc;
  if (precision == LENSIM_FP64)
    rc = run_linear<double>(m, g, psf, kx, ky, edge, ss, data, d_params, n_img, source_marg, linear_prior, d_amp,
                            d_logl, static_cast<double *>(d_model_out), d_cov,
                            static_cast<double *>(d_response_out), w, LL, st);
  else
    rc = run_linear<float>(m, g, psf, kx, ky, edge, ss, data, d_params, n_img, source_marg, linear_prior, d_amp,
                           d_logl, static_cast<float *>(d_model_out), d_cov, static_cast<float *>(d_response_out), w,
                           LL, st);
  return cr.end(st, rc);
}

size_t lensim_grad_workspace_size(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                                  int32_t n_img, int32_t precision) {
  ModelArgs m;
  if (model_args(model, m) || !grid) return 0;
  int kx = 1, ky = 1, edge = 0;
  if (psf_dims(psf, 0, kx, ky, edge)) return 0;
  GridArgs g;
  SSPlan ss;
  if (grid_args(grid, g) || psf_ss(psf, g, 0, ss)) return 0;
  return grad_layout(m, grid->nx, grid->ny, kx, ky, n_img, precision, edge, ss, g.s).total;
}

int lensim_loglike_grad(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                        const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                        double *d_logl, double *d_grad, void *ws, size_t ws_bytes, void *stream) {
  ModelArgs m;
  int rc = model_args(model, m);
  if (rc) return rc;
  GridArgs g;
  if ((rc = grid_args(grid, g))) return rc;
  int kx, ky, edge;
  if ((rc = psf_dims(psf, 0, kx, ky, edge))) return rc;
  SSPlan ss;
  if ((rc = psf_ss(psf, g, 0, ss))) return rc;
  if (!data || !data->d_data) return fail(LENSIM_EINVAL, "data image is NULL");
  if (!data->d_noise_map && !data->d_exposure_map && !(data->exposure_time > 0.0))
    return fail(LENSIM_EINVAL, "need a noise map, an exposure map or a positive exposure time");
  if (data->image_stride != 0 && data->image_stride < int64_t(g.nx) * g.ny)
    return fail(LENSIM_EINVAL, "image_stride %lld is smaller than an image", (long long)data->image_stride);
  if (n_img < 0) return fail(LENSIM_EINVAL, "negative n_img");
  if (n_img == 0) return LENSIM_OK;
  if (n_img > 65535 || m.n_prof > 65535) return fail(LENSIM_EINVAL, "batch too large");
  if (m.mp_n && int64_t(g.nx) * g.ny * g.s * g.s >= (int64_t(1) << 30))
    return fail(LENSIM_EINVAL, "multi-plane gradients take fewer than 2^30 rays per image");
  if (!d_logl || !d_grad || !d_params) return fail(LENSIM_EINVAL, "NULL buffer");
  if (precision != LENSIM_FP64 && precision != LENSIM_MIXED) return fail(LENSIM_EINVAL, "bad precision");
  const Layout L = grad_layout(m, g.nx, g.ny, kx, ky, n_img, precision, edge, ss, g.s);
  if ((rc = check_ws(L, ws, ws_bytes))) return rc;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  unsigned char *w = static_cast<unsigned char *>(ws);
  CallRanges cr;
  add_inputs(cr, model, m, d_params, n_img, grid, psf, data, g.nx, g.ny);
  cr.add(d_logl, size_t(n_img) * sizeof(double));
  cr.add(d_grad, size_t(n_img) * m.n_prof * LENSIM_NPARAM * sizeof(double));
  cr.add(ws, L.total);
  if ((rc = cr.begin(st))) return rc;
  if (precision == LENSIM_FP64)
    rc = run_grad<double>(m, g, psf, kx, ky, edge, ss, data, d_params, n_img, d_logl, d_grad, w, L, st);
  else
    rc = run_grad<float>(m, g, psf, kx, ky, edge, ss, data, d_params, n_img, d_logl, d_grad, w, L, st);
  return cr.end(st, rc);
}

size_t lensim_noalloc_scratch_bytes(void) { return kScratchBytes; }

int lensim_loglike_noalloc(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                           const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                           double *d_logl, void *d_model_out, void *stream) {
  if (n_img > 0) {
    const size_t need = lensim_workspace_size(model, grid, psf, n_img, precision);
    if (need == 0)  // a bad argument: lensim_loglike's own checks name it, before any device work
      return lensim_loglike(model, grid, psf, data, d_params, n_img, precision, d_logl, d_model_out, nullptr, 0,
                            stream);
    if (need > kScratchBytes)
      return fail(LENSIM_EWORKSPACE,
                  "the call needs %zu B of scratch memory; lensim_loglike_noalloc holds %zu B (split the batch, "
                  "or call lensim_loglike with a workspace)",
                  need, kScratchBytes);
  }
  void *ws = nullptr;
  const int rc = scratch_on_device(&ws);
  if (rc) return rc;
  return lensim_loglike(model, grid, psf, data, d_params, n_img, precision, d_logl, d_model_out, ws, kScratchBytes,
                        stream);
}

// ------------------------------------------------------ graph capture ---
struct lensim_graph_s {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

int lensim_graph_capture_loglike(const lensim_model *model, const lensim_grid *grid, const lensim_psf *psf,
                                 const lensim_data *data, const double *d_params, int32_t n_img, int32_t precision,
                                 double *d_logl, void *d_model_out, void *ws, size_t ws_bytes, void *stream,
                                 lensim_graph *out) {
  if (!out) return fail(LENSIM_EINVAL, "graph handle pointer is NULL");
  *out = nullptr;
  if (!stream) return fail(LENSIM_EINVAL, "graph capture needs a stream other than the NULL stream");
  if (n_img <= 0) return fail(LENSIM_EINVAL, "graph capture needs n_img > 0");
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
  g_capturing = true;
  const int rc = lensim_loglike(model, grid, psf, data, d_params, n_img, precision, d_logl, d_model_out, ws,
                                ws_bytes, stream);
  g_capturing = false;
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(st, &graph);  // ends the capture whatever the call returned
  if (rc != LENSIM_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return hip_fail(e, "hipGraphInstantiate");
  }
  lensim_graph g = static_cast<lensim_graph>(malloc(sizeof(lensim_graph_s)));
  if (!g) {
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    return fail(LENSIM_EINVAL, "out of host memory");
  }
  g->graph = graph;
  g->exec = exec;
  *out = g;
  return LENSIM_OK;
}

int lensim_graph_launch(lensim_graph g, void *stream) {
  if (!g) return fail(LENSIM_EINVAL, "graph is NULL");
  const hipError_t e = hipGraphLaunch(g->exec, reinterpret_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  return LENSIM_OK;
}

int lensim_graph_destroy(lensim_graph g) {
  if (!g) return LENSIM_OK;
  const hipError_t e1 = hipGraphExecDestroy(g->exec);
  const hipError_t e2 = hipGraphDestroy(g->graph);
  free(g);
  if (e1 != hipSuccess) return hip_fail(e1, "hipGraphExecDestroy");
  if (e2 != hipSuccess) return hip_fail(e2, "hipGraphDestroy");
  return LENSIM_OK;
}

}  // extern "C"
