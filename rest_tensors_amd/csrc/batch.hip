// Batched RI pipeline for many molecules of uneven size (the grouped
// workload: nbas drawn from [64, 384], naux = 3 nbas, nocc = nbas/8).
//
// Every stage of the FP64 pipeline is ONE launch covering all molecules:
//   1 density      D_m = C_occ C_occ^T          grouped GEMM (persistent)
//   2 gamma        gamma_m,P = <A_m,P, D_m>      TN path: sum_o B_m[o, o, P] after
//                                                stage 5; generic path: one block
//                                                per (m, 2 slices) over ri3ao
//   3 J            J_m = sum_P A_m,P gamma_m,P   grouped, one lane per packed (i, j)
//   4 ao2mo s1     X^T = A_P^T C_occ             grouped TN GEMM (direct-to-LDS)
//   5 ao2mo s2     B_P = (X^T)^T C               grouped TN GEMM
//   6 K partials   sum_P B_P^T B_P (upper tiles) grouped TN GEMM, batch-reducing
// Stages 4-6 run on the TN engine, which needs k-contiguous operands with a
// depth and leading dimension that are multiples of 16 and 16-byte aligned
// rows.  The batch keeps its own zero-padded layouts (nbp, noccp = nb, nocc
// rounded up to 16): ri3ao as Ap [nbp, nbp, naux] (copied once at creation,
// again by rt_batch_reload), C as Cp [nbp, nb], X^T as [nbp, naux, noccp]
// (rows (nu, P) = nu + nbp P) and B as Bp [ldbp, nb] (rows (P, o) = P + naux o
// for o < nocc, ldbp = naux nocc rounded up to 16, the rows past naux nocc
// zero).  Each stage is then one GEMM per molecule with the aux index inside
// the row index:
//   4  X^T[(nu,P), o]  = sum_mu Ap[mu, (nu,P)] Cp[mu, o]     M = nbp naux
//   5  Bp[(P,o), i]    = sum_nu X^T[nu, (P,o)] Cp[nu, i]      M = naux nocc
//   6  K slots         = sum over (P,o) of Bp[(P,o), i] Bp[(P,o), j], the
//                        depth (P,o) contiguous, split into 16-deep batches
//                        that the partial slots sum in order.
// The caller's B ([nocc, nb, naux]) is filled from Bp by a transposing copy.
// RTA_BATCH_GROUPED=1 keeps stages 4-6 on the generic grouped engine (no
// padded ri3ao; X^T per slice as [nbp, noccp]).
//   7 K reduce     slot sums + mirror            grouped, one block per (m, 32-tile)
//   8 energies     E_J, E_x per molecule         one block per molecule
// The eight launches are captured once into a HIP graph and replayed.  The
// grouped GEMMs walk a device work list (descriptor + prefix sums); their
// tile edge (64 or 128) and persistent grid size come from an autotuner that
// times every candidate on the actual molecule set and keeps the fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <string>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"
#include "mfma_f64.hpp"
#include "mfma_tn.hpp"
#include "timer.hpp"

namespace rta {

struct MolDev {
  const double* A;
  const double* C;
  double *D, *J, *K, *B;
  double *X, *gamma, *part;
  double *Bp, *Cp;  // padded B [naux, noccp, nb] and C [nbp, nb] (X^T: see above)
  double* Ap;       // padded ri3ao [nbp, nbp, naux] (TN path only)
  int64_t nb, naux, nocc;
  int64_t nbp, noccp;
  int64_t ldbp;     // rows of Bp: naux*nocc rounded up to 16 (zero rows at the end)
  int64_t nchunks;  // K partial slots
  int64_t bchunk;   // per slot: aux slices (generic path) or 16-deep batches (TN path)
};

__device__ __forceinline__ int find_mol(const int64_t* __restrict__ offs, int nmol,
                                        int64_t g) {
  int lo = 0, hi = nmol;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (offs[mid] <= g) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double b_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// one block per (molecule, pair of aux slices): each D element read serves two
// slices; upper triangle of A_P with weights 2 (off-diagonal) and 1
__global__ __launch_bounds__(256) void batch_gamma_kernel(
    const MolDev* __restrict__ mols, const int64_t* __restrict__ offs, int nmol) {
  __shared__ double red[2][4];
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t P = 2 * (blockIdx.x - offs[m]);
  const int64_t nb = M.nb, n2 = nb * nb;
  const bool two = P + 1 < M.naux;
  const double* Ap = M.A + P * n2;
  const double* Aq = two ? Ap + n2 : Ap;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double acc = 0.0, acc2 = 0.0;
  for (int64_t j = wave; j < nb; j += 4) {
    const double* cp = Ap + j * nb;
    const double* cq = Aq + j * nb;
    const double* dcol = M.D + j * nb;
    for (int64_t i = lane; i <= j; i += 64) {
      const double dw = (i == j) ? dcol[i] : 2.0 * dcol[i];
      acc += cp[i] * dw;
      acc2 += cq[i] * dw;
    }
  }
  acc = b_wave_sum(acc);
  acc2 = b_wave_sum(acc2);
  if (lane == 0) {
    red[0][wave] = acc;
    red[1][wave] = acc2;
  }
  __syncthreads();
  if (threadIdx.x == 0) M.gamma[P] = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
  if (threadIdx.x == 1 && two)
    M.gamma[P + 1] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
}

// The same sums from the padded copy Ap (leading dimension nbp, a multiple of
// 16, so every column starts 16-byte aligned), one block per (molecule, SP
// aux slices): a lane takes the pair of rows (2t, 2t+1) of a column with one
// TN path: with D = C_occ C_occ^T, gamma_P = <A_P, D> = sum_o B[o, o, P]
// (B_P = C_occ^T A_P C), so gamma is read off the occupied diagonal of Bp
// (row (P, o) = P + naux o, column i: B[o, i, P]) after ao2mo stage 2 instead
// of from another pass over ri3ao.  One lane per (molecule, P), the o-sum in
// order.
__global__ __launch_bounds__(256) void batch_gamma_bdiag_kernel(
    const MolDev* __restrict__ mols, const int64_t* __restrict__ offs, int nmol) {
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t P = (blockIdx.x - offs[m]) * 256 + threadIdx.x;
  if (P >= M.naux) return;
  const double* p = M.Bp + P;
  const int64_t step = M.naux + M.ldbp;  // (P, o) -> (P, o + 1), column o -> o + 1
  double acc = 0.0;
  for (int64_t o = 0; o < M.nocc; ++o) acc += p[o * step];
  M.gamma[P] = acc;
}

// one block per (molecule, 256 consecutive pairs of the packed upper
// triangle, q = j(j+1)/2 + i): every lane owns one pair (i, j) and sums all
// aux slices with 16 loads in flight, so short columns leave no lane idle;
// writes J[i,j] and J[j,i]
__global__ __launch_bounds__(256) void batch_j_kernel(const MolDev* __restrict__ mols,
                                                      const int64_t* __restrict__ offs,
                                                      int nmol) {
  constexpr int U = 16;
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t nb = M.nb, n2 = nb * nb, naux = M.naux;
  const int64_t q = (blockIdx.x - offs[m]) * 256 + threadIdx.x;
  if (q >= nb * (nb + 1) / 2) return;
  int64_t j = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while (j * (j + 1) / 2 > q) --j;
  while ((j + 1) * (j + 2) / 2 <= q) ++j;
  const int64_t i = q - j * (j + 1) / 2;
  const double* p = M.A + i + j * nb;
  const double* gm = M.gamma;
  double acc = 0.0, acc2 = 0.0;
  int64_t P = 0;
  for (; P + U - 1 < naux; P += U) {
    double a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = p[(P + u) * n2];
#pragma unroll
    for (int u = 0; u < U; u += 2) {
      acc += a[u] * gm[P + u];
      acc2 += a[u + 1] * gm[P + u + 1];
    }
  }
  for (; P < naux; ++P) acc += p[P * n2] * gm[P];
  const double v = acc + acc2;
  M.J[i + j * nb] = v;
  if (i != j) M.J[j + i * nb] = v;
}

// TN path: the same sums over the padded copy Ap (columns 16-byte aligned,
// ld nbp), one lane per pair of rows (i, i+1), i even, of column j's upper
// part with 16-byte loads; column j holds floor(j/2) + 1 row pairs, so the
// pairs of columns < j number j^2/4 rounded up.  Per element the sum runs in
// the order of batch_j_kernel (results bit-identical).
__device__ __forceinline__ int64_t bj_pairs_before(int64_t j) {
  return (j / 2) * (j / 2 + 1) + (j & 1) * (j / 2 + 1);
}

__global__ __launch_bounds__(256) void batch_j_x2_kernel(const MolDev* __restrict__ mols,
                                                         const int64_t* __restrict__ offs,
                                                         int nmol) {
  constexpr int U = 16;
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t nb = M.nb, nbp = M.nbp, n2h = nbp * nbp / 2, naux = M.naux;
  const int64_t q = (blockIdx.x - offs[m]) * 256 + threadIdx.x;
  if (q >= bj_pairs_before(nb)) return;
  int64_t j = (int64_t)(2.0 * sqrt((double)q));
  while (j > 0 && bj_pairs_before(j) > q) --j;
  while (bj_pairs_before(j + 1) <= q) ++j;
  const int64_t i = 2 * (q - bj_pairs_before(j));
  const double2* p = reinterpret_cast<const double2*>(M.Ap + i + j * nbp);
  const double* gm = M.gamma;
  double accx = 0.0, accx2 = 0.0, accy = 0.0, accy2 = 0.0;
  int64_t P = 0;
  for (; P + U - 1 < naux; P += U) {
    double2 a[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = p[(P + u) * n2h];
#pragma unroll
    for (int u = 0; u < U; u += 2) {
      accx += a[u].x * gm[P + u];
      accx2 += a[u + 1].x * gm[P + u + 1];
      accy += a[u].y * gm[P + u];
      accy2 += a[u + 1].y * gm[P + u + 1];
    }
  }
  for (; P < naux; ++P) {
    const double2 a = p[P * n2h];
    accx += a.x * gm[P];
    accy += a.y * gm[P];
  }
  const double vx = accx + accx2;
  M.J[i + j * nb] = vx;
  if (i != j) M.J[j + i * nb] = vx;
  if (i + 1 <= j) {
    const double vy = accy + accy2;
    M.J[(i + 1) + j * nb] = vy;
    if (i + 1 != j) M.J[j + (i + 1) * nb] = vy;
  }
}

// one block per (molecule, upper 32x32 tile): K = sum of partial slots, mirrored
__global__ __launch_bounds__(256) void batch_kreduce_kernel(
    const MolDev* __restrict__ mols, const int64_t* __restrict__ offs, int nmol) {
  __shared__ double T[32][33];
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  int tm, tn;
  upper_tile(blockIdx.x - offs[m], tm, tn);
  const int64_t nb = M.nb, slot = nb * nb;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t r = (int64_t)tm * 32 + tx, c = (int64_t)tn * 32 + ty + 8 * q;
    double s = 0.0;
    if (r < nb && c < nb && r <= c) {
      const double* p = M.part + r + c * nb;
      for (int64_t k = 0; k < M.nchunks; ++k) s += p[k * slot];
      M.K[r + c * nb] = s;
    }
    T[ty + 8 * q][tx] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t c = (int64_t)tn * 32 + tx, r = (int64_t)tm * 32 + ty + 8 * q;
    if (r < nb && c < nb && r < c) M.K[c + r * nb] = T[tx][ty + 8 * q];
  }
}

// one block per molecule: energies[2m] = E_J = 1/2 sum gamma^2,
// energies[2m+1] = E_x = -1/2 sum_{n<nocc} K[n,n]
// Cp(nu, i) = C(nu, i), nu < nb; one block per (molecule, column i).  Rows
// nb..nbp of Cp are zero from creation and never written.
__global__ __launch_bounds__(256) void batch_cpad_kernel(const MolDev* __restrict__ mols,
                                                         const int64_t* __restrict__ offs,
                                                         int nmol) {
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t i = blockIdx.x - offs[m];
  for (int64_t nu = threadIdx.x; nu < M.nb; nu += 256) M.Cp[nu + M.nbp * i] = M.C[nu + M.nb * i];
}

// Ap[mu, nu, P] = A[mu, nu, P] (zero outside nb x nb); one block per (molecule, P)
__global__ __launch_bounds__(256) void batch_apad_kernel(const MolDev* __restrict__ mols,
                                                         const int64_t* __restrict__ offs,
                                                         int nmol) {
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t P = blockIdx.x - offs[m];
  const int64_t nb = M.nb, nbp = M.nbp;
  const double* src = M.A + P * nb * nb;
  double* dst = M.Ap + P * nbp * nbp;
  for (int64_t e = threadIdx.x; e < nbp * nbp; e += 256) {
    const int64_t mu = e % nbp, nu = e / nbp;
    dst[e] = (mu < nb && nu < nb) ? src[mu + nb * nu] : 0.0;
  }
}

// B[o, i, P] = Bp[(P, o), i] for o < nocc: per column i a transpose between
// P-fastest (Bp) and o-fastest (B) orders.  One block per (molecule, column
// i, 32 aux slices): the 32 x nocc block goes through LDS, read along P and
// written along o (nocc <= kBcopyMaxOcc; larger nocc takes several passes).
constexpr int kBcopyMaxOcc = 64;

__global__ __launch_bounds__(256) void batch_bcopy_kernel(const MolDev* __restrict__ mols,
                                                          const int64_t* __restrict__ offs,
                                                          int nmol) {
  __shared__ double T[kBcopyMaxOcc][33];
  const int m = find_mol(offs, nmol, blockIdx.x);
  const MolDev M = mols[m];
  const int64_t loc = blockIdx.x - offs[m];
  const int64_t npt = (M.naux + 31) / 32;
  const int64_t i = loc / npt, P0 = (loc % npt) * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const double* src = M.Bp + M.ldbp * i;                  // (P, o) at P + naux o
  for (int64_t o0 = 0; o0 < M.nocc; o0 += kBcopyMaxOcc) {
    const int64_t oc = (M.nocc - o0 < kBcopyMaxOcc) ? M.nocc - o0 : kBcopyMaxOcc;
    __syncthreads();
    for (int64_t o = ty; o < oc; o += 8) {
      const int64_t P = P0 + tx;
      if (P < M.naux) T[o][tx] = src[P + M.naux * (o0 + o)];
    }
    __syncthreads();
    // write: for each P of the block, oc consecutive o (lanes along o)
    for (int idx = threadIdx.x; idx < 32 * oc; idx += 256) {
      const int pp = (int)(idx / oc);
      const int64_t o = idx % oc;
      const int64_t P = P0 + pp;
      if (P < M.naux) M.B[(o0 + o) + M.nocc * i + M.nocc * M.nb * P] = T[o][pp];
    }
  }
}

__global__ __launch_bounds__(256) void batch_energy_kernel(const MolDev* __restrict__ mols,
                                                           double* __restrict__ energies) {
  __shared__ double red[2][4];
  const MolDev M = mols[blockIdx.x];
  double a = 0.0, b = 0.0;
  for (int64_t P = threadIdx.x; P < M.naux; P += 256) a += M.gamma[P] * M.gamma[P];
  for (int64_t n = threadIdx.x; n < M.nocc; n += 256) b += M.K[n + n * M.nb];
  a = b_wave_sum(a);
  b = b_wave_sum(b);
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = a;
    red[1][threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    energies[2 * blockIdx.x] = 0.5 * ((red[0][0] + red[0][1]) + (red[0][2] + red[0][3]));
    energies[2 * blockIdx.x + 1] =
        -0.5 * ((red[1][0] + red[1][1]) + (red[1][2] + red[1][3]));
  }
}

// ---- host side ------------------------------------------------------------------
struct GroupCfg {
  int bm = 64, bn = 64;  // tile; the upper-tile K stage uses square tiles
  int wm = 2;            // waves along M (4 / wm along N)
  int grid_mul = 2;      // persistent workgroups per CU
};

// Tile candidates per stage.  Stages 1 and 2 have one side nocc = nbas/8 wide
// (8..48 for the grouped workload): narrow tiles on that side, the 4 waves
// stacked along the wide one.
static std::vector<GroupCfg> stage_candidates(int s) {
  if (s == 1) return {{64, 64, 2}, {128, 128, 2}, {128, 32, 4}, {128, 48, 4}, {256, 32, 4},
                      {256, 48, 4}};
  if (s == 2) return {{64, 64, 2}, {128, 128, 2}, {32, 128, 1}, {48, 128, 1}, {32, 256, 1},
                      {48, 256, 1}};
  return {{64, 64, 2}, {128, 128, 2}};
}

struct GroupedStage {
  bool ta = true, tb = false;
  GroupCfg cfg;
  std::vector<GemmParams> descs;
  GemmParams* d_descs = nullptr;
  int64_t* d_offs = nullptr;
  int64_t total = 0;
};

// Stages 5-6 on the grouped TN engine (NSTAGE 2, BK 16)
struct TnStage {
  GroupCfg cfg;
  std::vector<TnParams> descs;
  TnParams* d_descs = nullptr;
  int64_t* d_offs = nullptr;
  int64_t total = 0;
};

static std::vector<GroupCfg> tn_candidates(int s) {
  if (s == 1) return {{256, 48, 4}, {128, 48, 4}, {128, 64, 2}, {128, 32, 4}, {256, 32, 4},
                      {256, 64, 4}};
  if (s == 2) return {{128, 128, 2}, {64, 128, 2}, {64, 64, 2}, {32, 128, 1}};
  return {{64, 64, 2}, {128, 128, 2}};
}

struct Batch {
  int nmol = 0;
  std::vector<MolDev> mols;
  MolDev* d_mols = nullptr;
  int64_t *d_off_gamma = nullptr, *d_off_j = nullptr, *d_off_kred = nullptr;
  int64_t* d_off_j2 = nullptr;  // (molecule, 256 row pairs) for batch_j_x2_kernel
  int64_t n_j2 = 0;
  int64_t n_gamma = 0, n_j = 0, n_kred = 0;
  double* energies = nullptr;  // device [2*nmol]
  std::vector<void*> owned;    // device allocations of the handle
  GroupedStage st[4];          // density, ao2mo s1, ao2mo s2, K partials
  TnStage tn[4];               // TN forms of stages 2 and 3 (used when use_tn)
  bool use_tn = true;
  int64_t* d_off_cpad = nullptr;  // (molecule, column) offsets for the C padding
  int64_t n_cpad = 0;
  int64_t* d_off_bcopy = nullptr;  // (molecule, column, 32 slices) offsets for the B copy
  int64_t n_bcopy = 0;
  int64_t* d_off_gpair = nullptr;  // (molecule, pair of aux slices) offsets for gamma
  int64_t n_gpair = 0;
  int64_t* d_off_gdiag = nullptr;  // (molecule, 256 aux slices): gamma from Bp
  int64_t n_gdiag = 0;
  double tune_ms[4] = {0, 0, 0, 0};
  // the RI-J branch (density, gamma, J) runs on a second stream beside the
  // ao2mo/K branch; events fork it from and join it to the caller's stream
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int dalloc(Batch* b, size_t bytes, void** p) {
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(RT_ERR_NOMEM, "batch: allocation of %zu bytes failed", bytes);
  }
  b->owned.push_back(*p);
  return RT_OK;
}

static GemmParams make_params(const GemmDesc& d, int bm, int bn, bool reduce, int64_t bchunk) {
  GemmParams g{};
  g.A = d.A;
  g.B = d.B;
  g.C = d.C;
  g.M = d.M;
  g.N = d.N;
  g.K = d.K;
  g.lda = d.lda;
  g.ldb = d.ldb;
  g.sA = d.sA;
  g.sB = d.sB;
  g.c_rs = d.c_rs;
  g.c_cs = d.c_cs;
  g.sC = d.sC;
  g.alpha = d.alpha;
  g.beta = d.beta;
  g.nbatch = d.batch;
  g.tiles_m = (int)((d.M + bm - 1) / bm);
  g.tiles_n = (int)((d.N + bn - 1) / bn);
  if (reduce) {
    g.bchunk = bchunk;
    g.nchunks = (d.batch + bchunk - 1) / bchunk;
    g.ntiles = (int64_t)g.tiles_m * (g.tiles_m + 1) / 2;
    g.reduce = 1;
    g.upper = 1;
  } else {
    g.bchunk = 1;
    g.nchunks = d.batch;
    g.ntiles = (int64_t)g.tiles_m * g.tiles_n;
  }
  g.nwork = g.ntiles * g.nchunks;
  return g;
}

// Descriptors of stage s for tile configuration c.  For the K stage the batch
// chunking is fixed by the partial-slot layout chosen at creation.
static void build_stage(Batch* b, int s, const GroupCfg& c) {
  const int bm = c.bm, bn = c.bn;
  GroupedStage& S = b->st[s];
  S.descs.clear();
  for (const MolDev& m : b->mols) {
    GemmDesc d;
    d.alpha = 1.0;
    if (s == 0) {  // D = C_occ C_occ^T
      d.ta = false;
      d.tb = true;
      d.M = d.N = m.nb;
      d.K = m.nocc;
      d.A = m.C;
      d.lda = m.nb;
      d.B = m.C;
      d.ldb = m.nb;
      d.C = m.D;
      d.c_cs = m.nb;
      S.descs.push_back(make_params(d, bm, bn, false, 1));
    } else if (s == 1) {  // X^T_P = A_P^T C_occ
      d.M = m.nb;
      d.N = m.nocc;
      d.K = m.nb;
      d.A = m.A;
      d.lda = m.nb;
      d.sA = m.nb * m.nb;
      d.B = m.C;
      d.ldb = m.nb;
      d.C = m.X;  // padded [nbp, noccp] per slice
      d.c_cs = m.nbp;
      d.sC = m.nbp * m.noccp;
      d.batch = m.naux;
      S.descs.push_back(make_params(d, bm, bn, false, 1));
    } else if (s == 2) {  // B_P = (X^T_P)^T C
      d.M = m.nocc;
      d.N = m.nb;
      d.K = m.nb;
      d.A = m.X;
      d.lda = m.nbp;
      d.sA = m.nbp * m.noccp;
      d.B = m.C;
      d.ldb = m.nb;
      d.C = m.B;
      d.c_cs = m.nocc;
      d.sC = m.nocc * m.nb;
      d.batch = m.naux;
      S.descs.push_back(make_params(d, bm, bn, false, 1));
    } else {  // K partials: sum_P B_P^T B_P, upper tiles
      d.M = d.N = m.nb;
      d.K = m.nocc;
      d.A = m.B;
      d.lda = m.nocc;
      d.sA = m.nocc * m.nb;
      d.B = m.B;
      d.ldb = m.nocc;
      d.sB = m.nocc * m.nb;
      d.C = m.part;
      d.c_rs = 1;
      d.c_cs = m.nb;
      d.sC = m.nb * m.nb;
      d.batch = m.naux;
      GemmParams g = make_params(d, bm, bn, true, m.bchunk);
      S.descs.push_back(g);
    }
  }
  S.cfg.bm = bm;
  S.cfg.bn = bn;
  S.cfg.wm = c.wm;
  S.ta = (s != 0);
  S.tb = (s == 0);
}

static void build_tn_stage(Batch* b, int s, const GroupCfg& c) {
  TnStage& S = b->tn[s];
  S.descs.clear();
  for (const MolDev& m : b->mols) {
    TnParams p{};
    p.alpha = 1.0;
    p.beta = 0.0;
    p.nbatch = 1;
    p.bchunk = 1;
    p.nchunks = 1;
    if (s == 1) {  // X^T[(nu,P), o] = sum_mu Ap[mu, (nu,P)] Cp[mu, o], o < nocc
      p.A = m.Ap;
      p.lda = m.nbp;
      p.B = m.Cp;
      p.ldb = m.nbp;
      p.C = m.X;
      p.c_rs = 1;
      p.c_cs = m.nbp * m.naux;
      p.M = m.nbp * m.naux;
      p.N = m.nocc;  // columns nocc..noccp of X^T stay zero
      p.K = m.nbp;
    } else if (s == 2) {  // Bp[(P,o), i] = sum_nu X^T[nu, (P,o)] Cp[nu, i]
      p.A = m.X;          // row (P, o) = P + naux o at nbp (P + naux o)
      p.lda = m.nbp;
      p.B = m.Cp;
      p.ldb = m.nbp;
      p.C = m.Bp;
      p.c_rs = 1;
      p.c_cs = m.ldbp;
      p.M = m.naux * m.nocc;  // rows (P, o), o < nocc
      p.N = m.nb;
      p.K = m.nbp;
    } else {  // K slots: depth (P,o) of each row of Bp, 16-deep batches
      p.A = m.Bp;
      p.B = m.Bp;
      p.lda = p.ldb = m.ldbp;
      p.sA = p.sB = 16;
      p.C = m.part;
      p.c_rs = 1;
      p.c_cs = m.nb;
      p.sC = m.nb * m.nb;
      p.M = p.N = m.nb;
      p.K = 16;
      p.nbatch = m.ldbp / 16;
      p.bchunk = m.bchunk;
      p.nchunks = m.nchunks;
      p.reduce = 1;
      p.upper = 1;
    }
    p.tiles_m = (int)((p.M + c.bm - 1) / c.bm);
    p.tiles_n = (int)((p.N + c.bn - 1) / c.bn);
    p.ntiles = p.upper ? (int64_t)p.tiles_m * (p.tiles_m + 1) / 2
                       : (int64_t)p.tiles_m * p.tiles_n;
    p.nwork = p.ntiles * p.nchunks;
    S.descs.push_back(p);
  }
  S.cfg.bm = c.bm;
  S.cfg.bn = c.bn;
  S.cfg.wm = c.wm;
}

static int upload_tn_stage(Batch* b, int s) {
  TnStage& S = b->tn[s];
  std::vector<int64_t> offs(S.descs.size() + 1, 0);
  for (size_t i = 0; i < S.descs.size(); ++i) offs[i + 1] = offs[i] + S.descs[i].nwork;
  S.total = offs.back();
  if (!S.d_descs) {
    RTA_TRY(dalloc(b, sizeof(TnParams) * S.descs.size(), (void**)&S.d_descs));
    RTA_TRY(dalloc(b, sizeof(int64_t) * offs.size(), (void**)&S.d_offs));
  }
  RTA_HIP(hipMemcpy(S.d_descs, S.descs.data(), sizeof(TnParams) * S.descs.size(),
                    hipMemcpyHostToDevice));
  RTA_HIP(hipMemcpy(S.d_offs, offs.data(), sizeof(int64_t) * offs.size(),
                    hipMemcpyHostToDevice));
  return RT_OK;
}

template <int BM, int BN, int WM, bool CS>
static void launch_tn_t(const TnStage& S, unsigned grid, hipStream_t s) {
  constexpr size_t lds = 2 * (BM + BN) * 16 * sizeof(double);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void*)dgemm_grouped_tn_kernel<BM, BN, 16, 2, WM, CS>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    attr = true;
  }
  hipLaunchKernelGGL((dgemm_grouped_tn_kernel<BM, BN, 16, 2, WM, CS>), dim3(grid), dim3(256),
                     lds, s, S.d_descs, S.d_offs, (int)S.descs.size());
}

// Stage 1 (N = nocc of each molecule, below the tile width for most) runs the
// column-skipping form; stage 2 and the K slots fill their tiles.
static int launch_tn(const TnStage& S, hipStream_t s, bool colskip) {
  if (S.total == 0) return RT_OK;
  const int64_t want = (int64_t)device_cu_count() * S.cfg.grid_mul;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, S.total));
  const GroupCfg& c = S.cfg;
  if (colskip) {
    if (c.bm == 256 && c.bn == 48 && c.wm == 4) launch_tn_t<256, 48, 4, true>(S, grid, s);
    else if (c.bm == 128 && c.bn == 48 && c.wm == 4) launch_tn_t<128, 48, 4, true>(S, grid, s);
    else if (c.bm == 128 && c.bn == 64 && c.wm == 2) launch_tn_t<128, 64, 2, true>(S, grid, s);
    else if (c.bm == 128 && c.bn == 32 && c.wm == 4) launch_tn_t<128, 32, 4, true>(S, grid, s);
    else if (c.bm == 256 && c.bn == 32 && c.wm == 4) launch_tn_t<256, 32, 4, true>(S, grid, s);
    else if (c.bm == 256 && c.bn == 64 && c.wm == 4) launch_tn_t<256, 64, 4, true>(S, grid, s);
    else return fail(RT_ERR_ARG, "batch: no stage-1 TN kernel for tile %dx%d (%d waves along M)",
                     c.bm, c.bn, c.wm);
  } else if (c.bm == 32 && c.bn == 128 && c.wm == 1) launch_tn_t<32, 128, 1, false>(S, grid, s);
  else if (c.bm == 64 && c.bn == 128 && c.wm == 2) launch_tn_t<64, 128, 2, false>(S, grid, s);
  else if (c.bm == 64 && c.bn == 64 && c.wm == 2) launch_tn_t<64, 64, 2, false>(S, grid, s);
  else if (c.bm == 128 && c.bn == 128 && c.wm == 2) launch_tn_t<128, 128, 2, false>(S, grid, s);
  else return fail(RT_ERR_ARG, "batch: no TN kernel for tile %dx%d (%d waves along M)", c.bm, c.bn,
                   c.wm);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

static int upload_stage(Batch* b, int s) {
  GroupedStage& S = b->st[s];
  std::vector<int64_t> offs(S.descs.size() + 1, 0);
  for (size_t i = 0; i < S.descs.size(); ++i) offs[i + 1] = offs[i] + S.descs[i].nwork;
  S.total = offs.back();
  if (!S.d_descs) {
    RTA_TRY(dalloc(b, sizeof(GemmParams) * S.descs.size(), (void**)&S.d_descs));
    RTA_TRY(dalloc(b, sizeof(int64_t) * offs.size(), (void**)&S.d_offs));
  }
  RTA_HIP(hipMemcpy(S.d_descs, S.descs.data(), sizeof(GemmParams) * S.descs.size(),
                    hipMemcpyHostToDevice));
  RTA_HIP(hipMemcpy(S.d_offs, offs.data(), sizeof(int64_t) * offs.size(),
                    hipMemcpyHostToDevice));
  return RT_OK;
}

template <int BM, int BN, int WM, bool TA, bool TB>
static void launch_grouped_t(const GroupedStage& S, unsigned grid, hipStream_t s) {
  hipLaunchKernelGGL((dgemm_grouped_kernel<BM, BN, 16, TA, TB, WM>), dim3(grid), dim3(256), 0,
                     s, S.d_descs, S.d_offs, (int)S.descs.size());
}

template <bool TA, bool TB>
static int launch_grouped_cfg(const GroupedStage& S, unsigned grid, hipStream_t s) {
  const GroupCfg& c = S.cfg;
#define RTA_G(BM, BN, WM) \
  if (c.bm == BM && c.bn == BN && c.wm == WM) { launch_grouped_t<BM, BN, WM, TA, TB>(S, grid, s); return RT_OK; }
  RTA_G(64, 64, 2)
  RTA_G(128, 128, 2)
  if constexpr (TA && !TB) {
    RTA_G(128, 32, 4)
    RTA_G(128, 48, 4)
    RTA_G(256, 32, 4)
    RTA_G(256, 48, 4)
    RTA_G(32, 128, 1)
    RTA_G(48, 128, 1)
    RTA_G(32, 256, 1)
    RTA_G(48, 256, 1)
  }
#undef RTA_G
  return fail(RT_ERR_ARG, "batch: no kernel for tile %dx%d (%d waves along M)", c.bm, c.bn, c.wm);
}

static int launch_grouped(const GroupedStage& S, hipStream_t s) {
  const int64_t want = (int64_t)device_cu_count() * S.cfg.grid_mul;
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(want, S.total));
  if (S.total == 0) return RT_OK;
  if (S.ta) RTA_TRY((launch_grouped_cfg<true, false>(S, grid, s)));
  else RTA_TRY((launch_grouped_cfg<false, true>(S, grid, s)));
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

// RI-J (gamma -> J; J is one HBM-bound pass over the upper triangles of
// ri3ao) runs on the side stream beside the MFMA-bound rest of the ao2mo/K
// chain and is joined before the energies.  TN path: forked after ao2mo
// stage 2, gamma is read off Bp's occupied diagonal and J overlaps the B copy
// and the K slots.  Generic path: forked after stage 1, gamma is its own pass
// over ri3ao.  Under stream capture the fork and join become graph edges.
static int run_gamma_j(Batch* b, hipStream_t sj) {
  if (b->use_tn && b->n_gdiag > 0)
    hipLaunchKernelGGL(batch_gamma_bdiag_kernel, dim3((unsigned)b->n_gdiag), dim3(256), 0, sj,
                       b->d_mols, b->d_off_gdiag, b->nmol);
  else if (!b->use_tn && b->n_gpair > 0)
    hipLaunchKernelGGL(batch_gamma_kernel, dim3((unsigned)b->n_gpair), dim3(256), 0, sj,
                       b->d_mols, b->d_off_gpair, b->nmol);
  if (b->use_tn && b->n_j2 > 0)
    hipLaunchKernelGGL(batch_j_x2_kernel, dim3((unsigned)b->n_j2), dim3(256), 0, sj, b->d_mols,
                       b->d_off_j2, b->nmol);
  else if (b->n_j > 0)
    hipLaunchKernelGGL(batch_j_kernel, dim3((unsigned)b->n_j), dim3(256), 0, sj, b->d_mols,
                       b->d_off_j, b->nmol);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

static int run_stages(Batch* b, hipStream_t s) {
  RTA_TRY(launch_grouped(b->st[0], s));  // density
  auto fork = [&]() -> int {
    RTA_HIP(hipEventRecord(b->ev_fork, s));
    RTA_HIP(hipStreamWaitEvent(b->side, b->ev_fork, 0));
    RTA_TRY(run_gamma_j(b, b->side));
    if (b->use_tn && b->n_bcopy > 0)  // Bp -> caller's B, beside the K slots (both read Bp)
      hipLaunchKernelGGL(batch_bcopy_kernel, dim3((unsigned)b->n_bcopy), dim3(256), 0, b->side,
                         b->d_mols, b->d_off_bcopy, b->nmol);
    RTA_HIP(hipGetLastError());
    RTA_HIP(hipEventRecord(b->ev_join, b->side));
    return RT_OK;
  };
  if (b->use_tn) {
    if (b->n_cpad > 0)
      hipLaunchKernelGGL(batch_cpad_kernel, dim3((unsigned)b->n_cpad), dim3(256), 0, s,
                         b->d_mols, b->d_off_cpad, b->nmol);
    RTA_HIP(hipGetLastError());
    RTA_TRY(launch_tn(b->tn[1], s, true));   // ao2mo stage 1 -> X^T
    RTA_TRY(launch_tn(b->tn[2], s, false));  // ao2mo stage 2 -> Bp
    RTA_TRY(fork());
    RTA_TRY(launch_tn(b->tn[3], s, false));  // K partials from Bp
  } else {
    RTA_TRY(launch_grouped(b->st[1], s));  // ao2mo stage 1
    RTA_TRY(fork());
    RTA_TRY(launch_grouped(b->st[2], s));  // ao2mo stage 2
    RTA_TRY(launch_grouped(b->st[3], s));  // K partials
  }
  if (b->n_kred > 0)
    hipLaunchKernelGGL(batch_kreduce_kernel, dim3((unsigned)b->n_kred), dim3(256), 0, s,
                       b->d_mols, b->d_off_kred, b->nmol);
  RTA_HIP(hipStreamWaitEvent(s, b->ev_join, 0));  // gamma, J (and the B copy) are done
  hipLaunchKernelGGL(batch_energy_kernel, dim3(b->nmol), dim3(256), 0, s, b->d_mols,
                     b->energies);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

static int launch_stage(Batch* b, int s, hipStream_t stream) {
  if (b->use_tn && s >= 1) return launch_tn(b->tn[s], stream, s == 1);
  return launch_grouped(b->st[s], stream);
}

static double time_stage(Batch* b, int s, hipStream_t stream, int reps) {
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1e30;
  if (launch_stage(b, s, stream) != RT_OK) return 1e30;  // warm
  (void)hipEventRecord(e0, stream);
  for (int r = 0; r < reps; ++r)
    if (launch_stage(b, s, stream) != RT_OK) return 1e30;
  (void)hipEventRecord(e1, stream);
  (void)hipEventSynchronize(e1);
  float ms = 0.f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms / reps;
}

}  // namespace rta

using namespace rta;

extern "C" int rt_batch_create(const rt_mol_desc* mols, int nmol, void** handle) {
  RTA_CHECK(mols && handle && nmol > 0, "bad arguments");
  Batch* b = new Batch();
  b->nmol = nmol;
  auto fail_cleanup = [&](int st) {
    for (void* p : b->owned) (void)hipFree(p);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->side) (void)hipStreamDestroy(b->side);
    delete b;
    return st;
  };
  if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess)
    return fail_cleanup(fail(RT_ERR_HIP, "batch: stream or event creation failed"));
  // RTA_BATCH_GROUPED=1 keeps stages 4-6 on the generic grouped engine
  const char* forced = getenv("RTA_BATCH_GROUPED");
  const bool use_tn = !(forced && atoi(forced) == 1);
  b->use_tn = use_tn;
  int64_t ng = 0, nj = 0, nk = 0, nc = 0, nbc = 0, ngp = 0, ngd = 0, nj2 = 0;
  std::vector<int64_t> og(1, 0), oj(1, 0), ok(1, 0), oc(1, 0), obc(1, 0), ogp(1, 0),
      ogd(1, 0), oj2(1, 0);
  for (int i = 0; i < nmol; ++i) {
    const rt_mol_desc& m = mols[i];
    if (!(m.nbas > 0 && m.naux > 0 && m.nocc > 0 && m.nocc <= m.nbas && m.ri3ao && m.C &&
          m.D && m.J && m.K && m.B))
      return fail_cleanup(fail(RT_ERR_ARG, "molecule %d: bad shape or NULL buffer", i));
    MolDev d{};
    d.A = m.ri3ao;
    d.C = m.C;
    d.D = m.D;
    d.J = m.J;
    d.K = m.K;
    d.B = m.B;
    d.nb = m.nbas;
    d.naux = m.naux;
    d.nocc = m.nocc;
    d.nbp = (m.nbas + 15) / 16 * 16;
    d.noccp = (m.nocc + 15) / 16 * 16;
    d.ldbp = (m.naux * m.nocc + 15) / 16 * 16;
    // K partial slots: up to 8, every slot non-empty (the reduce sums them
    // all); the TN path slices the depth (P, o) in 16-deep batches, the
    // generic one the aux slices
    const int64_t nbt = use_tn ? d.ldbp / 16 : m.naux;
    const int64_t want = std::max<int64_t>(1, std::min<int64_t>(nbt, 8));
    d.bchunk = (nbt + want - 1) / want;
    d.nchunks = (nbt + d.bchunk - 1) / d.bchunk;
    const size_t xbytes = sizeof(double) * d.nbp * d.noccp * m.naux;
    const size_t cbytes = sizeof(double) * d.nbp * m.nbas;
    int st = dalloc(b, xbytes, (void**)&d.X);
    if (st == RT_OK) st = dalloc(b, sizeof(double) * m.naux, (void**)&d.gamma);
    if (st == RT_OK)
      st = dalloc(b, sizeof(double) * d.nchunks * m.nbas * m.nbas, (void**)&d.part);
    const size_t bpbytes = sizeof(double) * d.ldbp * m.nbas;
    if (st == RT_OK) st = dalloc(b, bpbytes, (void**)&d.Bp);
    if (st == RT_OK) st = dalloc(b, cbytes, (void**)&d.Cp);
    if (st == RT_OK && use_tn)
      st = dalloc(b, sizeof(double) * d.nbp * d.nbp * m.naux, (void**)&d.Ap);
    // the padding of X^T, Cp and Bp is zero for the life of the handle (never written)
    if (st == RT_OK && (hipMemset(d.X, 0, xbytes) != hipSuccess ||
                        hipMemset(d.Cp, 0, cbytes) != hipSuccess ||
                        hipMemset(d.Bp, 0, bpbytes) != hipSuccess))
      st = fail(RT_ERR_HIP, "batch: clearing the padded buffers failed");
    if (st != RT_OK) return fail_cleanup(st);
    b->mols.push_back(d);
    const int64_t t32 = (m.nbas + 31) / 32;
    ng += m.naux;
    nj += (m.nbas * (m.nbas + 1) / 2 + 255) / 256;
    {
      const int64_t h = m.nbas / 2, npairs = h * (h + 1) + (m.nbas & 1) * (h + 1);
      nj2 += (npairs + 255) / 256;
      oj2.push_back(nj2);
    }
    nk += t32 * (t32 + 1) / 2;
    nc += m.nbas;
    nbc += m.nbas * ((m.naux + 31) / 32);
    obc.push_back(nbc);
    ngp += (m.naux + 1) / 2;
    ogp.push_back(ngp);
    ngd += (m.naux + 255) / 256;
    ogd.push_back(ngd);
    og.push_back(ng);
    oj.push_back(nj);
    ok.push_back(nk);
    oc.push_back(nc);
  }
  b->n_gamma = ng;
  b->n_j = nj;
  b->n_j2 = nj2;
  b->n_kred = nk;
  b->n_cpad = nc;
  b->n_bcopy = nbc;
  b->n_gpair = ngp;
  b->n_gdiag = ngd;

  int st = dalloc(b, sizeof(MolDev) * nmol, (void**)&b->d_mols);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * og.size(), (void**)&b->d_off_gamma);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * oj.size(), (void**)&b->d_off_j);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * oj2.size(), (void**)&b->d_off_j2);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * ok.size(), (void**)&b->d_off_kred);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * oc.size(), (void**)&b->d_off_cpad);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * obc.size(), (void**)&b->d_off_bcopy);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * ogp.size(), (void**)&b->d_off_gpair);
  if (st == RT_OK) st = dalloc(b, sizeof(int64_t) * ogd.size(), (void**)&b->d_off_gdiag);
  if (st == RT_OK) st = dalloc(b, sizeof(double) * 2 * nmol, (void**)&b->energies);
  if (st != RT_OK) return fail_cleanup(st);
  if (hipMemcpy(b->d_mols, b->mols.data(), sizeof(MolDev) * nmol, hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_off_gamma, og.data(), sizeof(int64_t) * og.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_off_j, oj.data(), sizeof(int64_t) * oj.size(), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(b->d_off_j2, oj2.data(), sizeof(int64_t) * oj2.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_off_kred, ok.data(), sizeof(int64_t) * ok.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_off_cpad, oc.data(), sizeof(int64_t) * oc.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_off_bcopy, obc.data(), sizeof(int64_t) * obc.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_off_gpair, ogp.data(), sizeof(int64_t) * ogp.size(),
                hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_off_gdiag, ogd.data(), sizeof(int64_t) * ogd.size(),
                hipMemcpyHostToDevice) != hipSuccess)
    return fail_cleanup(fail(RT_ERR_HIP, "batch: descriptor upload failed"));
  for (int s = 0; s < 4; ++s) {
    build_stage(b, s, GroupCfg{});
    st = upload_stage(b, s);
    if (st != RT_OK) return fail_cleanup(st);
  }
  for (int s = 1; s < 4; ++s) {
    build_tn_stage(b, s, tn_candidates(s)[0]);
    st = upload_tn_stage(b, s);
    if (st != RT_OK) return fail_cleanup(st);
  }
  if (use_tn) {  // the padded copy of ri3ao
    hipLaunchKernelGGL(batch_apad_kernel, dim3((unsigned)b->n_gamma), dim3(256), 0, nullptr,
                       b->d_mols, b->d_off_gamma, b->nmol);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
      return fail_cleanup(fail(RT_ERR_HIP, "batch: padded ri3ao copy failed"));
  }
  *handle = b;
  return RT_OK;
}

extern "C" int rt_batch_autotune(void* h, int reps, void* stream) {
  RTA_CHECK(h, "NULL handle");
  Batch* b = (Batch*)h;
  hipStream_t s = (hipStream_t)stream;
  reps = std::max(1, reps);
  // stages 1 and 2 read D / X written by earlier stages: run once so every
  // candidate sees valid inputs (results are recomputed by rt_batch_run)
  RTA_TRY(run_stages(b, s));
  RTA_HIP(hipStreamSynchronize(s));
  for (int stg = 0; stg < 4; ++stg) {
    if (b->use_tn && stg >= 1) {
      double best = 1e30;
      GroupCfg bestc;
      for (const GroupCfg& cand : tn_candidates(stg)) {
        build_tn_stage(b, stg, cand);
        RTA_TRY(upload_tn_stage(b, stg));
        for (int gm : {1, 2, 4}) {
          b->tn[stg].cfg.grid_mul = gm;
          const double ms = time_stage(b, stg, s, reps);
          if (ms < best) {
            best = ms;
            bestc = b->tn[stg].cfg;
          }
        }
      }
      build_tn_stage(b, stg, bestc);
      RTA_TRY(upload_tn_stage(b, stg));
      b->tn[stg].cfg = bestc;
      b->tune_ms[stg] = best;
      continue;
    }
    double best = 1e30;
    GroupCfg bestc;
    for (const GroupCfg& cand : stage_candidates(stg)) {
      build_stage(b, stg, cand);
      RTA_TRY(upload_stage(b, stg));
      for (int gm : {1, 2, 4}) {
        b->st[stg].cfg.grid_mul = gm;
        const double ms = time_stage(b, stg, s, reps);
        if (ms < best) {
          best = ms;
          bestc = b->st[stg].cfg;
        }
      }
    }
    build_stage(b, stg, bestc);
    RTA_TRY(upload_stage(b, stg));
    b->st[stg].cfg = bestc;
    b->tune_ms[stg] = best;
  }
  if (b->exec) {  // the graph baked in the old descriptors' launch shapes
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  return RT_OK;
}

extern "C" int rt_batch_set_stage(void* h, int stage, int bm, int bn, int waves_m,
                                  int grid_mul) {
  RTA_CHECK(h, "NULL handle");
  RTA_CHECK(stage >= 0 && stage < 4, "stage must be 0..3 (density, ao2mo s1, ao2mo s2, K)");
  RTA_CHECK(grid_mul >= 1 && grid_mul <= 64, "grid_mul must be 1..64");
  Batch* b = (Batch*)h;
  const bool tn = b->use_tn && stage >= 1;
  bool known = false;
  for (const GroupCfg& c : tn ? tn_candidates(stage) : stage_candidates(stage))
    known = known || (c.bm == bm && c.bn == bn && c.wm == waves_m);
  RTA_CHECK(known, "stage %d has no %dx%d tile with %d waves along M", stage, bm, bn, waves_m);
  GroupCfg c;
  c.bm = bm;
  c.bn = bn;
  c.wm = waves_m;
  if (tn) {
    build_tn_stage(b, stage, c);
    RTA_TRY(upload_tn_stage(b, stage));
    b->tn[stage].cfg.grid_mul = grid_mul;
  } else {
    build_stage(b, stage, c);
    RTA_TRY(upload_stage(b, stage));
    b->st[stage].cfg.grid_mul = grid_mul;
  }
  if (b->exec) {  // a captured graph holds the old launch shapes
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  return RT_OK;
}

extern "C" int rt_batch_reload(void* h, void* stream) {
  RTA_CHECK(h, "NULL handle");
  Batch* b = (Batch*)h;
  if (!b->use_tn || b->n_gamma == 0) return RT_OK;
  hipLaunchKernelGGL(batch_apad_kernel, dim3((unsigned)b->n_gamma), dim3(256), 0,
                     (hipStream_t)stream, b->d_mols, b->d_off_gamma, b->nmol);
  RTA_HIP(hipGetLastError());
  return RT_OK;
}

extern "C" int rt_batch_run(void* h, void* stream) {
  RTA_CHECK(h, "NULL handle");
  KernelTimer kt("batch_run", (hipStream_t)stream);
  return run_stages((Batch*)h, (hipStream_t)stream);
}

extern "C" int rt_batch_graph_capture(void* h, void* stream) {
  RTA_CHECK(h && stream, "graph capture needs a handle and a non-null stream");
  Batch* b = (Batch*)h;
  hipStream_t s = (hipStream_t)stream;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
  }
  // run_stages records no timing events, so the graph holds kernel nodes only
  RTA_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int st = run_stages(b, s);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (st != RT_OK) {
    if (g) (void)hipGraphDestroy(g);
    return st;
  }
  if (e != hipSuccess)
    return fail(RT_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  RTA_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph = g;
  return RT_OK;
}

extern "C" int rt_batch_graph_launch(void* h, void* stream) {
  RTA_CHECK(h, "NULL handle");
  Batch* b = (Batch*)h;
  RTA_CHECK(b->exec, "no captured graph (call rt_batch_graph_capture first)");
  KernelTimer kt("batch_graph", (hipStream_t)stream);
  RTA_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return RT_OK;
}

extern "C" int rt_batch_energies(void* h, double* e_j, double* e_x, void* stream) {
  RTA_CHECK(h && e_j && e_x, "NULL argument");
  Batch* b = (Batch*)h;
  std::vector<double> e(2 * b->nmol);
  RTA_HIP(hipMemcpyAsync(e.data(), b->energies, sizeof(double) * e.size(),
                         hipMemcpyDeviceToHost, (hipStream_t)stream));
  RTA_HIP(hipStreamSynchronize((hipStream_t)stream));
  for (int i = 0; i < b->nmol; ++i) {
    e_j[i] = e[2 * i];
    e_x[i] = e[2 * i + 1];
  }
  return RT_OK;
}

extern "C" int rt_batch_info(void* h, char* buf, int len) {
  RTA_CHECK(h && buf && len > 0, "NULL argument");
  Batch* b = (Batch*)h;
  static const char* names[4] = {"density", "ao2mo_s1", "ao2mo_s2", "k_partials"};
  std::string s = "{";
  for (int i = 0; i < 4; ++i) {
    char t[160];
    const bool tn = b->use_tn && i >= 1;
    const GroupCfg& c = tn ? b->tn[i].cfg : b->st[i].cfg;
    snprintf(t, sizeof t,
             "%s\"%s\": {\"engine\": \"%s\", \"tile\": [%d, %d], \"waves_m\": %d, "
             "\"grid_mul\": %d, \"tuned_ms\": %.4f}",
             i ? ", " : "", names[i], tn ? "tn" : "generic", c.bm, c.bn, c.wm, c.grid_mul,
             b->tune_ms[i]);
    s += t;
  }
  s += "}";
  snprintf(buf, len, "%s", s.c_str());
  return RT_OK;
}

extern "C" int rt_batch_destroy(void* h) {
  if (!h) return RT_OK;
  Batch* b = (Batch*)h;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  (void)hipDeviceSynchronize();  // the side stream's work is done before its buffers go
  for (void* p : b->owned) (void)hipFree(p);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->side) (void)hipStreamDestroy(b->side);
  delete b;
  return RT_OK;
}
