// otm_kernels.hpp -- host-side launchers of the gfx950 kernels.
#pragma once

#include <vector>

#include "otm_common.hpp"

namespace otm {

// ---- fine level (matrix-free, otm_fine.hip) ----
void launch_fine_gs_color(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* f,
                          float* u, int color);
void launch_fine_residual(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* f,
                          const float* u, float* res);
void launch_fine_apply_dot(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* p,
                           float* q, const Reduce& R, unsigned nblocks);
// x as an FP32 pair when xlo is given (nullptr: x alone)
void launch_fine_init_residual(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa,
                               const float* x, const float* xlo, float* r, const Reduce& R, unsigned nblocks);
void launch_fine_init_residual(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa,
                               const double* x, float* r, const Reduce& R, unsigned nblocks);
struct ChiRef;  // fluctuation fields as FP64, an FP32 pair or FP32 (otm_fine_dev.hpp)
// S6 (may be null): the element energies are also stored there (6 x nn floats)
void launch_fine_kappa_h(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const ChiRef& chi,
                         const Reduce& R, unsigned nblocks, float* S6 = nullptr);
// sens (pitched) = sum_c w_c S6[c] from stored element energies
void launch_sens_from_energy(hipStream_t s, const Geom& g, const float* S6, const float w[6], float* sens, int pitch,
                             const float* scale = nullptr);
// scale (may be null): each element's sensitivity times scale[e]
void launch_fine_sensitivity(hipStream_t s, const Geom& g, const FineConsts& fc, const ChiRef& chi, const float w[6],
                             float* sens, int pitch, const float* scale = nullptr);
void launch_fine_energy6(hipStream_t s, const Geom& g, const FineConsts& fc, const ChiRef& chi, float* S6);
// Fine brick smoother: tiles of bx x by x bz nodes, one node per thread and
// colour step for tr thread rows of bx threads, halo frozen, result in
// uout. Each tile shape is its own block Gauss-Seidel (the couplings that
// leave a tile are frozen during a smoothing step), so the shape is a
// per-context choice (otm_config.fine_tile, otm_autotune); variant 0 is the
// default where the cell tiles with it, else variant 1.
struct FineTileShape {
  int bx, by, bz, tr;
};
constexpr int kFineTileVariants = 5;
constexpr FineTileShape kFineTileShapes[kFineTileVariants] = {
    {32, 4, 8, 8}, {16, 4, 8, 8}, {32, 8, 4, 8}, {32, 4, 4, 4}, {32, 8, 8, 8}};
bool fine_tile_ok(const Geom& g, int variant);
int fine_tile_default(const Geom& g);  // -1: the cell does not tile
// fine_brick_x: bx of the default variant (0 when the cell does not tile).
// With dot (fine_brick_dot_ok), the blocks also store their partial sums of
// f.uout for that reducer (the PCG's r.z after a preconditioner
// application); launch_fine_brick_dot_final then finishes the reduction.
int fine_brick_x(const Geom& g);
bool fine_brick_dot_ok(const Geom& g, int variant);
void launch_fine_smooth_brick(hipStream_t s, const Geom& g, const FineConsts& fc, const float* kappa, const float* f,
                              const float* uin, float* uout, int sweeps, bool forward, bool zero_in, int variant,
                              const Reduce* dot = nullptr);
void launch_fine_brick_dot_final(hipStream_t s, const Geom& g, const Reduce& R, int variant);

// ---- coarse levels (stored 27-point stencils, otm_coarse.hip) ----
// Stored stencil of a coarse level: FP32 (f16 == nullptr) or FP16 couplings
// scaled by a per-level factor held on the device.
struct StRef {
  const float* f32;
  const _Float16* f16;
  const float* scale;
};
void launch_stencil_pack(hipStream_t s, const Geom& g, const float* st, float* scale, _Float16* st16);
void launch_st_gs_color(hipStream_t s, const Geom& g, const StRef& S, const float* f, float* u, int color);
void launch_st_residual(hipStream_t s, const Geom& g, const StRef& S, const float* f, const float* u, float* res);
void launch_restrict(hipStream_t s, const Geom& fine, const Geom& coarse, const float* res, float* fc);
void launch_prolong_add(hipStream_t s, const Geom& fine, const Geom& coarse, const float* uc, float* uf);
void launch_galerkin_first(hipStream_t s, const Geom& fine, const Geom& coarse, const GalerkinConsts& G,
                           const float* kappa, float* E);
void launch_galerkin_next(hipStream_t s, const Geom& fine, const Geom& coarse, const float* Ein, float* Eout);
void launch_stencil_from_elements(hipStream_t s, const Geom& g, const float* E, float* st);
// Device doubles the coarsest factorisation needs as work space for n nodes.
size_t coarsest_work_doubles(long long n);
void launch_coarsest_blocked(hipStream_t s, const Geom& g, const float* st, double* B, double* Binv);
void launch_coarsest_factor(hipStream_t s, const Geom& g, const float* st, double* B, double* Binv);
void launch_coarsest_solve(hipStream_t s, const Geom& g, const double* Binv, const float* f, float* u);

// Brick smoother (8^3 bricks, frozen halo, result in uout).
bool brick_ok(const Geom& g);
void launch_st_smooth_brick(hipStream_t s, const Geom& g, const StRef& S, const float* f, const float* uin,
                            float* uout, int sweeps, bool forward, bool zero_in);

// ---- vectors and reductions (otm_vec.hip) ----
// One-block second stage of a reduction over nblocks partial rows; above
// kMaxReduceBlocks rows (only for OTM_NRHS values: the smoother's per-tile
// r.z) the rows are summed in a loop.
void launch_reduce_final(hipStream_t s, const Reduce& R, int nv, unsigned nblocks);
void launch_fill(hipStream_t s, float* a, long long n, float v);
void launch_dot3(hipStream_t s, long long nn, const float* a, const float* b, const Reduce& R, unsigned nblocks);
void launch_sum3(hipStream_t s, long long nn, const float* a, const Reduce& R, unsigned nblocks);
void launch_sum3(hipStream_t s, long long nn, const double* a, const Reduce& R, unsigned nblocks);
void launch_refine_add(hipStream_t s, long long n, double* x64, float* d);
// x += d (x as an FP32 pair when xlo is given), then d = 0
void launch_add_pair(hipStream_t s, long long n, float* x, float* xlo, float* d);
void launch_sub_mean_cast(hipStream_t s, long long nn, double* x64, float* x32, const PcgScalars* P);
void launch_sub_mean3(hipStream_t s, long long nn, float* a, const PcgScalars* P);
void launch_sum2(hipStream_t s, long long nn, const float* a, const float* b, const Reduce& R, unsigned nblocks);
void launch_pcg_xr(hipStream_t s, long long nn, float* x, float* xlo, float* r, const float* p, const float* q,
                   const PcgScalars* P, const Reduce& R, unsigned nblocks);
// p = z + beta p, or p = z when first (a PCG start or restart)
void launch_pcg_p(hipStream_t s, long long nn, float* p, const float* z, const PcgScalars* P, bool first = false);
void launch_copy_pitched(hipStream_t s, const Geom& g, const float* src, int pitch, float* dst, float add);
void launch_copy_to_pitched(hipStream_t s, const Geom& g, const float* src, float* dst, int pitch);

// ---- density filter, SIMP and OC (otm_opt.hip) ----
constexpr int kMaxTaps = 343;  // (2*3+1)^3: radius up to 3.5 voxels
struct Taps {
  int n;                 // number of non-zero taps
  int periodic;          // wrap around the cell instead of renormalising
  signed char dx[kMaxTaps], dy[kMaxTaps], dz[kMaxTaps];
  float w[kMaxTaps];
};
// W: per-node weight sums, written (w_known false) or, when a previous call
// with the same taps wrote them, read for the nodes near a non-periodic
// cell's faces instead of being summed again.
void launch_filter_fwd(hipStream_t s, const Geom& g, const Taps& T, const float* in, float* out, float* W,
                       bool w_known = false);
void launch_filter_adj(hipStream_t s, const Geom& g, const Taps& T, const float* in, const float* W, float* out);
void launch_simp_fwd(hipStream_t s, long long n, const float* rho, float k0, float k1, float floor_, float* kappa,
                     float* dkappa);
void launch_mul(hipStream_t s, long long n, const float* a, const float* b, float* out);
void launch_oc_filter_sens(hipStream_t s, const Geom& g, const Taps& T, const float* sens, const float* rho,
                           float* out);
// Device-side OC bisection state (OCOptimizer::update, OCOptimizer.cu:38-67).
struct OcState {
  float lo, hi;   // bracket of the Lagrange multiplier (FP32 as in the reference)
  float g;        // last evaluated midpoint
  int done;       // volume within tolerance
  int steps;      // midpoints evaluated
  int pad;
  double vol;     // volume fraction at the last midpoint
};
void launch_maxabs(hipStream_t s, long long n, const float* a, unsigned int* out_bits);
// seeded cosine-mode field for the random initialisers
struct InitModes {
  int count;
  float mean;
  float k[16][3];
  float amp[16];
  float phase[16];
};
void launch_init_density(hipStream_t s, const Geom& g, int kind, const InitModes& M, float* rho);
// randTri (Framework.cuh:238-310): basis sum per node (period P <= 10),
// level-set map sums and application, and the noise initialiser's pieces
struct RandTriRot {
  float m[9];  // row-major rotation of the centred coordinates
};
int rand_tri_max_period();
void launch_rand_tri(hipStream_t s, const Geom& g, int P, int centered, const RandTriRot& R, const float* coeffs,
                     float* out);
unsigned level_sum_blocks(long long n);
void launch_level_sum(hipStream_t s, long long n, const float* val, float t, float hi, double* part);
void launch_level_apply(hipStream_t s, long long n, float* val, float t, float hi);
void launch_noise(hipStream_t s, long long n, unsigned long long seed, float* out);
void launch_project_binary(hipStream_t s, long long n, float* v);
void launch_scale_clamp(hipStream_t s, long long n, float* v, float scale, float lo, float hi);
void launch_oc_begin(hipStream_t s, long long n, const float* sens, unsigned int* bits, OcState* st);
constexpr int kOcTree = 5;                   // bisection levels per pass over the design
constexpr int kOcCand = (1 << kOcTree) - 1;  // midpoints evaluated per pass
// One pass: the candidate sums, then the one-block final stage that walks
// the tree (a second launch).
void launch_oc_tree_pass(hipStream_t s, long long n, const float* sens, const float* rho, OcState* st,
                         float min_rho, float step, float damp, double volratio, double vol_tol, int max_steps,
                         const Reduce& R, unsigned nblocks);
// The OC state as published to the host by k_oc_apply (coherent pinned memory).
struct OcHost {
  OcState st;
  int applied;    // the update was applied
  unsigned seq;   // tag of the apply launch that published this
};
// Applies the update at the state's multiplier when the bisection is done,
// has used max_steps midpoints, or force is set; otherwise leaves rho. Then
// publishes the state and whether it applied to hm with `tag`.
void launch_oc_apply(hipStream_t s, long long n, const float* sens, float* rho, const OcState* st, float min_rho,
                     float step, float damp, int max_steps, int force, OcHost* hm, unsigned tag);

// ---- MMA (otm_mma.hip) ----
constexpr int kMmaMaxM = 4;  // constraints supported by the device subsolver
constexpr int kMmaMaxNS = 2 * kMmaMaxM + (kMmaMaxM * (kMmaMaxM + 1)) / 2;  // Newton-system sums
struct MmaArrays {
  long long n;
  float *xvar, *xold1, *xold2;  // design and the two previous designs (FP32, as the caller holds them)
  const float *xmin, *xmax;     // per-element bounds, or null: xmin_u / xmax_u for every element
  double xmin_u, xmax_u, move;
  double *low, *upp;              // asymptotes, carried between updates
  const float *df0dx, *dgdx;      // gradients (dgdx: m x n)
  double *xs, *xsi, *eta;         // interior point
  double *xs_t, *xsi_t, *eta_t;   // its trial
  double* partials;               // [blocks * 32]
};
constexpr int kMmaMaxBlocks = 2048;  // partial-sum slots of the MMA passes
size_t mma_small_bytes();
// HIP-event timing of the MMA passes (otm_mma_profile_*): class 0 the
// line search's later trial-point passes, 1 each Newton step's first trial
// (which also forms the elements' step bound), 2 the point pass at the
// start of a barrier level; bytes are the algorithmic bytes per launch.
struct MmaProf {
  bool on = false;
  long long launches[3] = {0, 0, 0};
  double ms[3] = {0, 0, 0}, bytes[3] = {0, 0, 0};
  struct Pending {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
};
void mma_prof_release(MmaProf& P);
// What the host loop of an MMA update needs after each point pass, in
// coherent pinned host memory: the pass's finaliser writes it and then its
// tag, and the host polls the tag (no copy, no stream synchronisation).
struct MmaHost {
  double residumax, rmaxnew, rmaxnext;
  int accept, respec, fault;
  unsigned seq;
};
// One MMA update on device arrays; returns Newton steps or a negative code
// (-1 HIP error, -2 non-finite values, -3 unsupported m). mh is the host
// block (host address), mh_dev its device address.
int mma_update(hipStream_t s, int m, MmaArrays& A, void* small_dev, MmaHost* mh, MmaHost* mh_dev,
               const double* gval_dev, int itn, double a0, double av, double cv, double dv, float* xout,
               MmaProf* prof = nullptr);
void launch_f2d(hipStream_t s, long long n, const float* in, double* out);
void launch_fill_d(hipStream_t s, double* a, long long n, double v);

// Symmetry group acting on a periodic lattice: element i maps coordinate c
// to c' with c'[j] = c[perm[i][j]], reflected (n - 1 - c'[j]) where bit j of
// flip[i] is set. Permutations need a cubic lattice.
struct SymGroup {
  int n;
  signed char perm[48][3];
  unsigned char flip[48];
};
void launch_symmetrize(hipStream_t s, const Geom& g, const SymGroup& G, const float* in, float* out);

// Full reductions of a field (otm_field_reduce): block partials of the max
// and of sum / sum x^p / sum exp(x - shift) in item order, and the gradient.
constexpr int kReduceSum = 0, kReducePnorm = 1, kReduceLogSumExp = 2;
void launch_field_max(hipStream_t s, long long n, const float* x, double* part, unsigned nb);
void launch_field_sum(hipStream_t s, int kind, long long n, const float* x, double p, double shift, double* part,
                      unsigned nb);
void launch_field_grad(hipStream_t s, int kind, long long n, const float* x, double p, double value, double upstream,
                       float* g);

}  // namespace otm
