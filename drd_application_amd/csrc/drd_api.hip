// drd_api.hip -- host runtime and C-ABI (include/drd_gpu.h) of the gfx950
// implicit-solve path.
//
// One drd_ctx per grid instance owns every device buffer: the module-`global`
// fields, the Krylov workspace, DCT plans and eigenvalue tables.  Solves run
// stream-ordered on the context's stream; BiCGSTAB iterations are queued in
// chunks without host round trips, the device-side state machine
// (kfinalize) deciding when the kernels turn into no-ops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "kernels.hpp"
#include "unet.hpp"

using namespace drd;

static thread_local std::string g_err;

static int set_err(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

void drd::api_set_error(const std::string& msg) { g_err = msg; }

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return set_err(DRD_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));   \
  } while (0)

#define RCHK(expr)            \
  do {                        \
    int r_ = (expr);          \
    if (r_ != 0) return r_;   \
  } while (0)

// Learned correction of the velocity preconditioner: the pair store, the two
// networks (ny x nx for the preconditioner, crop size for training) and the
// judge/fallback state.
constexpr int kSnapMax = 16;        // recorded iterates per collected solve
constexpr double kLearnSlack = 1.25;  // allowed iteration growth over the judged plain count

struct Learn {
  UNet* net = nullptr;
  UNet* tr = nullptr;
  int ch = 0, cw = 0;
  int cap = 0, count = 0, next = 0;
  int64_t npx = 0;
  // crops > 0: a pair is one crop_h x crop_w crop (crops of them taken at
  // random from each recorded iterate); 0: a pair is the whole ny x nx image
  int crops = 0;
  int64_t spx = 0;           // pixels of a stored pair (npx, or crop_h * crop_w)
  uint16_t* feat = nullptr;  // cap x spx x 8 bf16
  float* targ = nullptr;     // cap x spx x 2
  uint16_t* fimg = nullptr;  // crops > 0: the iterate's whole feature image
  float* ftarg = nullptr;    //            and its target
  int every = 0;             // FGMRES directions j with j % every == 0 take the correction (0: j == 0)
  uint16_t* img = nullptr;   // npx x 8 (preconditioner input)
  float* out = nullptr;      // npx x 2
  uint16_t* cfeat = nullptr; // crop input
  float* ctarg = nullptr;
  double *ss = nullptr, *ss_part = nullptr;
  double *r = nullptr, *d = nullptr, *z = nullptr, *e = nullptr, *w = nullptr;
  double* dac = nullptr;  // [d | A c], contiguous for one two-vector reduction
  double *part = nullptr, *dots = nullptr;
  double* snaps = nullptr;   // kSnapMax x Nuv
  uint64_t rng = 0;
  int train_iters = 0;
  int state = DRD_LEARN_OFF;
  int ref_its = 0, judged_its = 0, last_its = 0, fallbacks = 0;
  double judged_ms[2] = {0, 0};  // learned FGMRES, plain method
  bool record = false;
  double loss_first = 0, loss_last = 0, train_ms = 0;
  std::vector<void*> allocs;
};

static void learn_destroy(Learn* L) {
  if (!L) return;
  unet_destroy(L->net);
  unet_destroy(L->tr);
  for (void* p : L->allocs) (void)hipFree(p);
  delete L;
}

// A captured chunk of BiCGSTAB iterations is valid for one operator,
// preconditioner, chunk length and pair of vectors; everything else its
// kernels bake in (workspace and coefficient pointers, the preconditioner
// objects and their buffers, the scalars, the grid and the batch group) is
// summed up in a fingerprint taken at capture and compared before each
// replay: a graph whose fingerprint no longer matches is captured again.
// Parameter changes and preconditioner rebuilds also drop every graph.
struct GraphKey {
  int op, prec, chunk;
  const double* b;
  const double* x;
  bool operator<(const GraphKey& o) const {
    return std::tie(op, prec, chunk, b, x) < std::tie(o.op, o.prec, o.chunk, o.b, o.x);
  }
};
struct GraphEntry {
  hipGraphExec_t exec = nullptr;
  uint64_t print = 0;
};

struct drd_ctx {
  drd_params p;
  Grid g;
  Scal s;
  hipStream_t stream = nullptr;
  // batch group: nb instances; instance i's per-instance buffers sit i * bs
  // bytes after instance 0's (one slab of bs bytes per instance)
  int nb = 1;
  int64_t bs = 0;
  void* slab = nullptr;
  // batch groups using multigrid: each instance's hierarchies (BDF2, BDF1)
  // inside its block of the slab
  void *mg_mem = nullptr, *mg1_mem = nullptr;
  std::vector<drd_solve_info> last;  // per instance: the latest BiCGSTAB solve
  std::vector<double> last_div;      // per instance: the latest mean divergence
  // captured BiCGSTAB iteration chunks (drd_use_graphs)
  bool use_graphs = false;
  int guess = 0;  // drd_set_initial_guess
  bool mg_fused = true;  // drd_mg_fused
  std::map<GraphKey, GraphEntry> graphs;
  int64_t graph_recaptures = 0;  // replays refused for a changed fingerprint
  int pred_its[5] = {0, 0, 0, 0, 0};  // per operator: iterations of its latest BiCGSTAB solve
  double phori_cb = 1.0, phori_ct = 1.0;  // PhSolve_ori's wall coefficients (coef_bot, coef_top of its last step)
  // drd_kernel_timer: event pairs around the instrumented kernel's launches
  int tm_which = -1, tm_n = 0;
  std::vector<hipEvent_t> tm_ev;
  double* f[DRD_F_COUNT] = {};
  // per-step scratch fields
  double *u_after = nullptr, *v_after = nullptr, *ph_after = nullptr, *rmu_after = nullptr;
  double* zero_c = nullptr;               // zero cell field (particle forces)
  double* arg[8] = {};                    // upload slots for host-array arguments
  // Krylov workspace
  int64_t nmax = 0;
  int64_t tiles = 0;  // kTile tiles of the longest vector
  double* kw[7] = {};                     // R RT P V T PH SH  (LIB/BiCGSTABREVCOM.f:159-166)
  double *rhs = nullptr, *xv = nullptr, *diag = nullptr, *tmp = nullptr, *tmp2 = nullptr;
  double* part = nullptr;
  KState* ks = nullptr;
  KState* ks_host = nullptr;              // pinned
  // tables
  double *pre_lcc = nullptr, *ph_lcc = nullptr, *cfB = nullptr, *cfT = nullptr;
  double* uL = nullptr;
  double* uL_host = nullptr;              // pinned
  double* pn = nullptr;
  DctPlan* dct = nullptr;
  VeloPrec* vprec = nullptr;  // fast-path velocity preconditioner (lazy)
  VeloMG* vmg = nullptr;      // multigrid velocity preconditioner (lazy)
  bool vprec_fp32 = true;
  // the same two for the BDF1 operator (lazy); vprec1 shares vprec's
  // per-instance buffers (one solve uses one of them at a time)
  VeloPrec* vprec1 = nullptr;
  bool vprec1_fp32 = true;
  VeloMG* vmg1 = nullptr;
  // FGMRES workspace (lazy): V (m+1 vectors), Z (m vectors), small arrays
  int fg_m = 0;
  double *fg_V = nullptr, *fg_Z = nullptr, *fg_h = nullptr, *fg_hh = nullptr;
  double* fg_part = nullptr;  // (m + 2) * 2 partial sums per tile
  // PhSolve / VCOST / ESOLP (lazy)
  Dct1Tab *d1_nx = nullptr, *d1_ny = nullptr, *d1_nx1 = nullptr;
  double *ph_f2 = nullptr, *ph_t = nullptr, *aux_work = nullptr, *esolp_tabs = nullptr;
  double* dct_work = nullptr;
  double* vp_inv64 = nullptr;  // per-instance buffers of the velocity preconditioner
  float* vp_inv32 = nullptr;
  void* vp_work = nullptr;
  void* rb_work = nullptr;    // red-black sweeps: the iterate between sweeps
  float* rb_c32 = nullptr;    // FP32 bu, M_inner, M_0 and block factors d11, d22, 1/det (nc each)
  double cosa = 0, sqrt2 = 0;
  Learn* learn = nullptr;  // drd_learn_init
  // drd_bicgstab_cb: the caller's MATVEC / PSOLVE, its vector length, and
  // pinned host copies of an operand and a result (nmax each, lazy)
  drd_matvec_cb ext_mv = nullptr;
  drd_psolve_cb ext_ps = nullptr;
  int64_t ext_n = 0;
  double *ext_h0 = nullptr, *ext_h1 = nullptr;
};

// operator / preconditioner ids of drd_bicgstab_cb (internal: a caller's own
// MATVEC and PSOLVE, called on the host)
constexpr int kOpExt = 100, kPsExt = 100;

BatchCfg& drd::batch_cfg() {
  static thread_local BatchCfg cfg;
  return cfg;
}

// Launches inside the scope cover every instance of the context's group.
struct BatchScope {
  BatchCfg saved;
  explicit BatchScope(const drd_ctx* c) : saved(batch_cfg()) { batch_cfg() = BatchCfg{c->nb, c->bs}; }
  ~BatchScope() { batch_cfg() = saved; }
};

// ---------------------------------------------------------------------------
extern "C" void drd_params_default(drd_params* p, int32_t nx, int32_t ny) {
  std::memset(p, 0, sizeof(*p));
  p->nx = nx;
  p->ny = ny;
  p->variant = DRD_VARIANT_TP;
  p->phi_mode = 2;
  p->lr_bd_type = 'n';
  p->velo_prec = DRD_PREC_REFERENCE;
  p->maxit_velo1 = 800;
  p->maxit_velo2 = 2000;
  p->maxit_phmu = 5000;
  p->maxit_ph = 500;
  p->dx = 1.0 / nx;
  p->dz = 1.0 / ny;
  p->dt = 1.0e-4 / 2.0;
  p->re = 3;
  p->ca = 12;
  p->xld = 5.0e-2;
  p->s_implicit = 1.5;
  p->I_thickness = 0.01;
  p->tol = 1.0e-8;
  p->penalty_k = 1.0e8;
  p->angle = 60;
  p->lambda_rho = 1;
  p->bdr_ph_used_adjust = 1.0e-6;
  p->twall = 0.1;
  p->bdr_upp = 1;
}

extern "C" const char* drd_last_error(void) { return g_err.c_str(); }
extern "C" int drd_abi_version(void) { return DRD_ABI_VERSION; }

extern "C" int64_t drd_field_size(const drd_params* p, int32_t f) {
  if (!p) return 0;
  const int64_t nc = (int64_t)(p->nx + 2) * (p->ny + 2);
  switch (f) {
    case DRD_F_U:
    case DRD_F_U_PRE:
      return (int64_t)(p->nx + 3) * (p->ny + 2);
    case DRD_F_V:
    case DRD_F_V_PRE:
      return (int64_t)(p->nx + 2) * (p->ny + 1);
    case DRD_F_ANGLEB:
    case DRD_F_ANGLET:
      return p->nx + 1;
    default:
      return (f >= 0 && f < DRD_F_COUNT) ? nc : 0;
  }
}

static Scal make_scal(const drd_params& p) {
  Scal s{};
  s.dt = p.dt;
  s.re = p.re;
  s.ca = p.ca;
  s.xld = p.xld;
  s.s = p.s_implicit;
  s.I = p.I_thickness;
  s.dr = p.dx;
  s.dz = p.dz;
  s.dr2 = p.dx * p.dx;
  s.dz2 = p.dz * p.dz;
  s.penalty_k = p.penalty_k;
  s.ggx = p.ggx;
  s.ggz = p.ggz;
  s.uwT = p.uwT;
  s.uwB = p.uwB;
  s.adj = p.bdr_ph_used_adjust;
  // coef_bot = (I/dz)/(I/dz)  TP/ph_mu_solve.f90:122-123
  const double tmp = p.I_thickness / p.dz;
  s.coef_bot = (p.I_thickness / p.dz) / tmp;
  s.coef_top = (p.I_thickness / p.dz) / tmp;
  s.lr_d = (p.lr_bd_type == 'd') || p.variant == DRD_VARIANT_MF;
  s.mf = p.variant == DRD_VARIANT_MF;
  s.phi_mode = p.phi_mode;
  scal_derive(s);
  return s;
}

static const double kPiG = 3.14159265358979323846264338327950288419716939937510;  // TP/global.f90:39

template <class T>
static int dalloc(T** ptr, int64_t n) {
  HIPCHK(hipMalloc((void**)ptr, (size_t)std::max<int64_t>(n, 1) * sizeof(T)));
  HIPCHK(hipMemset(*ptr, 0, (size_t)std::max<int64_t>(n, 1) * sizeof(T)));
  return 0;
}

// Ph_Lcc TP/setup.f90:137-155 for the step dt, into the context's table
static int build_ph_lcc(drd_ctx* c, double dt) {
  const drd_params& p = c->p;
  const int nx = p.nx, ny = p.ny;
  const double dr2 = p.dx * p.dx, dz2 = p.dz * p.dz;
  std::vector<double> fmx(nx + 1), fmy(ny + 1), ph((size_t)nx * ny);
  for (int i = 0; i < nx; ++i) {
    const double sn = std::sin(kPiG * i / 2.0 / (nx - 1));
    fmx[i] = 4.0 * (sn * sn) / dr2;
  }
  for (int j = 0; j < ny; ++j) {
    const double sn = std::sin(kPiG * j / 2.0 / (ny - 1));
    fmy[j] = 4.0 * (sn * sn) / dz2;
  }
  for (int j = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i) {
      const double temp = fmx[i] + fmy[j];
      ph[i + (size_t)nx * j] =
          1. + p.xld * dt * p.s_implicit * temp / p.I_thickness + p.xld * p.I_thickness * dt * (temp * temp);
    }
  HIPCHK(hipMemcpy(c->ph_lcc, ph.data(), ph.size() * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

static int build_tables(drd_ctx* c) {
  const drd_params& p = c->p;
  const int nx = p.nx, ny = p.ny;
  const double dr2 = p.dx * p.dx, dz2 = p.dz * p.dz;
  // Pre_Lcc and Ph_Lcc  TP/setup.f90:108-155
  std::vector<double> fmx(nx + 1), fmy(ny + 1), pre((size_t)(nx + 1) * (ny + 1)), ph((size_t)nx * ny);
  for (int i = 0; i <= nx; ++i) {
    const double sn = std::sin(kPiG * i / 2.0 / nx);
    fmx[i] = 4.0 * (sn * sn) / dr2;
  }
  for (int j = 0; j <= ny; ++j) {
    const double sn = std::sin(kPiG * j / 2.0 / ny);
    fmy[j] = 4.0 * (sn * sn) / dz2;
  }
  if (p.lr_bd_type == 'p') {
    // periodic x: 4 sin^2(pi kx / nx) / dr2 of the Fourier mode kx at the
    // halfcomplex indices m = kx (cosine part) and nx - kx (sine part)
    for (int m = 0; m < nx; ++m) {
      const int kx = m <= nx / 2 ? m : nx - m;
      const double sn = std::sin(kPiG * kx / nx);
      fmx[m] = 4.0 * (sn * sn) / dr2;
    }
    fmx[nx] = 0.0;
  }
  for (int j = 0; j <= ny; ++j)
    for (int i = 0; i <= nx; ++i) pre[i + (size_t)(nx + 1) * j] = -(fmx[i] + fmy[j]);
  pre[0] = -1.0;
  RCHK(dalloc(&c->pre_lcc, (int64_t)pre.size()));
  RCHK(dalloc(&c->ph_lcc, (int64_t)ph.size()));
  HIPCHK(hipMemcpy(c->pre_lcc, pre.data(), pre.size() * sizeof(double), hipMemcpyHostToDevice));
  RCHK(build_ph_lcc(c, p.dt));
  // ESOLP tables TP/setup.f90:157-235 (BDF1 set).  t1 keeps the value of the
  // last loop iteration (:167, :203), (pi*nx/nx)/2 and (pi*(nx-1)/(nx-1))/2,
  // overwriting the mass shift of :159 as the reference does.
  {
    const int64_t nu_t = (int64_t)ny * (nx + 1), nv_t = (int64_t)(ny - 1) * nx, nt = nu_t + nv_t;
    std::vector<double> ta(nt), tb(nt), tc(nt), fx(nx + 1);
    const double xb = (2.0 * 0.0 - p.dz) / (2.0 * 0.0 + p.dz);  // no-slip: xlsB = xlsT = 0
    double t1 = 0.0;
    for (int i = 0; i <= nx; ++i) {
      t1 = kPiG * i / nx / 2.;
      const double sn = std::sin(t1);
      fx[i] = 4.0 * p.dt * (sn * sn) / dr2;
    }
    for (int k = 0; k <= nx; ++k)
      for (int i = 1; i <= ny; ++i) {
        const int64_t q = (i - 1) + (int64_t)ny * k;
        if (i == 1) {
          ta[q] = 0.0;
          tb[q] = t1 + ((2.0 - xb) * p.dt / dz2 + fx[k]);
          tc[q] = (-1.0 * p.dt / dz2);
        } else if (i == ny) {
          ta[q] = (-1.0 * p.dt / dz2);
          tb[q] = t1 + ((2.0 - xb) * p.dt / dz2 + fx[k]);
          tc[q] = 0.0;
        } else {
          ta[q] = (-p.dt / dz2);
          tb[q] = t1 + (2.0 * p.dt / dz2 + fx[k]);
          tc[q] = (-p.dt / dz2);
        }
      }
    for (int i = 0; i < nx; ++i) {
      t1 = kPiG * i / (nx - 1) / 2.;
      const double sn = std::sin(t1);
      fx[i] = 4.0 * p.dt * (sn * sn) / dr2;
    }
    for (int k = 0; k < nx; ++k)
      for (int i = 1; i <= ny - 1; ++i) {
        const int64_t q = nu_t + (i - 1) + (int64_t)(ny - 1) * k;
        ta[q] = (i == 1) ? 0.0 : -p.dt / dz2;
        tb[q] = t1 + 2.0 * p.dt / dz2 + fx[k];
        tc[q] = (i == ny - 1) ? 0.0 : -p.dt / dz2;
        if (i == 1) tc[q] = -1.0 * p.dt / dz2;
        if (i == ny - 1) ta[q] = -1.0 * p.dt / dz2;
      }
    hipFree(c->esolp_tabs);
    c->esolp_tabs = nullptr;
    RCHK(dalloc(&c->esolp_tabs, 3 * nt));
    HIPCHK(hipMemcpy(c->esolp_tabs, ta.data(), nt * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->esolp_tabs + nt, tb.data(), nt * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->esolp_tabs + 2 * nt, tc.data(), nt * sizeof(double), hipMemcpyHostToDevice));
  }
  c->cosa = std::cos(p.angle * kPiG / 180);
  c->sqrt2 = std::sqrt(2.0);
  return 0;
}

// cf = cos(angle*pi/180)/sqrt(2) per column  TP/ch_bdr_condition.f90:61
// wetting coefficients of instance `inst` from its contact angles
static int upload_angles(drd_ctx* c, int inst = 0) {
  const int nx = c->p.nx;
  const int64_t o = (int64_t)inst * c->bs;
  auto at = [&](double* p) { return reinterpret_cast<double*>(reinterpret_cast<char*>(p) + o); };
  std::vector<double> aB(nx + 1), aT(nx + 1), cB(nx + 1), cT(nx + 1);
  HIPCHK(hipMemcpy(aB.data(), at(c->f[DRD_F_ANGLEB]), (nx + 1) * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(aT.data(), at(c->f[DRD_F_ANGLET]), (nx + 1) * sizeof(double), hipMemcpyDeviceToHost));
  for (int i = 0; i <= nx; ++i) {
    cB[i] = std::cos(aB[i] * kPiG / 180.0) / std::sqrt(2.0);
    cT[i] = std::cos(aT[i] * kPiG / 180.0) / std::sqrt(2.0);
  }
  HIPCHK(hipMemcpy(at(c->cfB), cB.data(), (nx + 1) * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(at(c->cfT), cT.data(), (nx + 1) * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

// Per-instance buffers come from one slab per instance, laid out alike;
// tables, upload slots and the lazily built FGMRES / PhSolve / ESOLP
// resources are separate allocations shared by the group.
static int ctx_alloc(drd_ctx* c) {
  const drd_params& p = c->p;
  const Grid& g = c->g;
  c->nmax = std::max<int64_t>(g.nuv(), 2 * g.ncell());
  const int64_t tiles = (c->nmax + kTile - 1) / kTile;
  if (tiles > kMaxTiles) return set_err(DRD_E_UNSUPPORTED, "grid too large for the reduction tree");
  c->tiles = tiles;
  struct Item {
    void** ptr;
    size_t bytes;
  };
  std::vector<Item> items;
  auto add = [&](auto** ptr, int64_t n, size_t esz) { items.push_back({(void**)ptr, (size_t)std::max<int64_t>(n, 1) * esz}); };
  for (int f = 0; f < DRD_F_COUNT; ++f) add(&c->f[f], drd_field_size(&p, f), 8);
  add(&c->u_after, g.nuf(), 8);
  add(&c->v_after, g.nvf(), 8);
  add(&c->ph_after, g.nc(), 8);
  add(&c->rmu_after, g.nc(), 8);
  add(&c->zero_c, g.nc(), 8);
  for (auto& w : c->kw) add(&w, c->nmax, 8);
  add(&c->rhs, c->nmax, 8);
  add(&c->xv, c->nmax, 8);
  add(&c->diag, c->nmax, 8);
  add(&c->tmp, c->nmax, 8);
  add(&c->tmp2, c->nmax, 8);
  // partial sums: up to 2 segments x 2 values per tile, or one per grid row
  add(&c->part, std::max<int64_t>(4 * tiles, g.ny + g.nx + 64), 8);
  add(&c->ks, 1, sizeof(KState));
  add(&c->pn, 2 * (int64_t)p.nx, 8);
  add(&c->dct_work, 2 * g.ncell(), 8);
  add(&c->vp_inv64, g.nuv(), 8);
  add(&c->vp_inv32, g.nuv(), 4);
  add(&c->vp_work, 2 * g.nuv(), 8);
  add(&c->rb_work, 3 * g.ncell(), 8);  // 2 ncell doubles, or 6 ncell floats
  add(&c->rb_c32, 6 * g.nc(), 4);
  add(&c->cfB, p.nx + 1, 8);
  add(&c->cfT, p.nx + 1, 8);
  if (c->nb > 1 && (p.velo_prec == DRD_PREC_MG || p.velo_prec == DRD_PREC_MG32)) {
    const size_t mgb = std::max(velo_mg_bytes(g, true), velo_mg_bytes(g, false));
    add(&c->mg_mem, (int64_t)mgb, 1);
    add(&c->mg1_mem, (int64_t)mgb, 1);
  }
  size_t S = 0;
  std::vector<size_t> off;
  for (const Item& it : items) {
    off.push_back(S);
    S += (it.bytes + 255) / 256 * 256;
  }
  c->bs = (int64_t)S;
  HIPCHK(hipMalloc(&c->slab, S * (size_t)c->nb));
  HIPCHK(hipMemset(c->slab, 0, S * (size_t)c->nb));
  for (size_t i = 0; i < items.size(); ++i) *items[i].ptr = static_cast<char*>(c->slab) + off[i];
  c->last.assign(c->nb, drd_solve_info{});
  c->last_div.assign(c->nb, 0.0);
  HIPCHK(hipHostMalloc((void**)&c->ks_host, c->nb * sizeof(KState), hipHostMallocDefault));
  const int64_t big = std::max<int64_t>(g.nuf(), g.nc());
  for (auto& a : c->arg) RCHK(dalloc(&a, big));
  RCHK(dalloc(&c->uL, p.ny + 1));
  HIPCHK(hipHostMalloc((void**)&c->uL_host, (p.ny + 1) * sizeof(double), hipHostMallocDefault));
  c->dct = dct_plan_create(p.nx, p.ny);
  return 0;
}

static void drop_graphs(drd_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second.exec);
  c->graphs.clear();
}

static void ctx_free(drd_ctx* c) {
  drop_graphs(c);
  if (c->ext_h0) hipHostFree(c->ext_h0);
  if (c->ext_h1) hipHostFree(c->ext_h1);
  c->ext_h0 = c->ext_h1 = nullptr;
  for (auto e : c->tm_ev) hipEventDestroy(e);
  c->tm_ev.clear();
  hipFree(c->slab);
  for (auto& a : c->arg) hipFree(a);
  if (c->ks_host) hipHostFree(c->ks_host);
  hipFree(c->pre_lcc);
  hipFree(c->ph_lcc);
  hipFree(c->uL);
  if (c->uL_host) hipHostFree(c->uL_host);
  dct_plan_destroy(c->dct);
  velo_prec_destroy(c->vprec);
  velo_mg_destroy(c->vmg);
  velo_prec_destroy(c->vprec1);
  velo_mg_destroy(c->vmg1);
  learn_destroy(c->learn);
  c->learn = nullptr;
  hipFree(c->fg_V);
  hipFree(c->fg_Z);
  hipFree(c->fg_h);
  hipFree(c->fg_part);
  if (c->fg_hh) hipHostFree(c->fg_hh);
  dct1_table_destroy(c->d1_nx);
  dct1_table_destroy(c->d1_ny);
  dct1_table_destroy(c->d1_nx1);
  hipFree(c->ph_f2);
  hipFree(c->ph_t);
  hipFree(c->aux_work);
  hipFree(c->esolp_tabs);
  if (c->stream) hipStreamDestroy(c->stream);
}

static VeloPrec* make_vprec(drd_ctx* c, bool fp32, int* err) {
  const VeloPrecBufs bufs{c->vp_inv64, c->vp_inv32, c->vp_work};
  return velo_prec_create(c->g, c->s, fp32, err, &bufs);
}

// The BDF1 operator (TP/VeloSolve_combined.f90:608-670) is the BDF2 stencil
// with mass coefficient re rho instead of 1.5 re rho and no penalty: the
// fast-path preconditioners built for BDF2 become BDF1 ones with these scalars.
static Scal bdf1_scal(const Scal& s) {
  Scal t = s;
  t.re = s.re / 1.5;
  t.penalty_k = 0.0;
  scal_derive(t);
  return t;
}

// The fast-Poisson velocity preconditioner is built with the context (or
// when drd_set_params first asks for it), so no solve pays its set-up.
static int eager_vprec(drd_ctx* c) {
  const int vp = c->p.velo_prec;
  if (vp != DRD_PREC_DCT32 && vp != DRD_PREC_DCT64 && vp != DRD_PREC_LEARNED) return 0;
  const bool fp32 = vp != DRD_PREC_DCT64;
  if (c->vprec && c->vprec_fp32 == fp32) return 0;
  drop_graphs(c);
  velo_prec_destroy(c->vprec);
  c->vprec = nullptr;
  int err = 0;
  c->vprec = make_vprec(c, fp32, &err);
  if (!c->vprec)
    return set_err(err ? err : DRD_E_HIP, "DCT velocity preconditioner needs power-of-two nx, ny <= 4096");
  c->vprec_fp32 = fp32;
  return 0;
}

// Periodic x (lr_bd_type 'p', no TP counterpart): the TP variant, nx a power
// of two (the pressure's real FFT in x), one instance, and the
// preconditioners built for it: the reference's (Jacobi / identity velocity,
// Jacobi phi/mu), Jacobi, identity and the phi/mu block solve.
static bool periodic_prec_ok(int prec) {
  return prec == DRD_PS_IDENTITY || prec == DRD_PS_VELO1_JACOBI || prec == DRD_PS_VELO2_JACOBI ||
         prec == DRD_PS_PHMU_JACOBI || prec == DRD_PS_PHMU_BLOCK;
}
static int check_periodic(const drd_params& p, int nb) {
  if (p.lr_bd_type != 'p') return 0;
  if (p.variant != DRD_VARIANT_TP) return set_err(DRD_E_UNSUPPORTED, "periodic x takes the TP variant");
  if ((p.nx & (p.nx - 1)) != 0) return set_err(DRD_E_UNSUPPORTED, "periodic x needs a power-of-two nx");
  if (nb != 1) return set_err(DRD_E_UNSUPPORTED, "periodic x takes a single-instance context");
  if (p.velo_prec != DRD_PREC_REFERENCE && p.velo_prec != DRD_PREC_IDENTITY && p.velo_prec != DRD_PREC_JACOBI)
    return set_err(DRD_E_UNSUPPORTED, "periodic x takes the reference, identity or Jacobi velocity preconditioner");
  if (p.ch_prec != DRD_CH_PREC_REFERENCE && p.ch_prec != DRD_CH_PREC_BLOCK)
    return set_err(DRD_E_UNSUPPORTED, "periodic x takes the reference or block phi/mu preconditioner");
  return 0;
}

static int ctx_create(const drd_params* prm, int nb, drd_ctx** out) {
  if (!prm || !out) return set_err(DRD_E_ARG, "null argument");
  if (prm->nx < 4 || prm->ny < 4) return set_err(DRD_E_ARG, "nx, ny must be >= 4");
  if (prm->nx > 4096 || prm->ny > 4096) return set_err(DRD_E_ARG, "nx, ny must be <= 4096");
  if (prm->lr_bd_type != 'n' && prm->lr_bd_type != 'd' && prm->lr_bd_type != 'p')
    return set_err(DRD_E_ARG, "lr_bd_type must be 'n', 'd' or 'p'");
  if (prm->phi_mode != 1 && prm->phi_mode != 2) return set_err(DRD_E_ARG, "phi_mode must be 1 or 2");
  if (prm->variant != DRD_VARIANT_TP && prm->variant != DRD_VARIANT_MF)
    return set_err(DRD_E_ARG, "variant must be DRD_VARIANT_TP or DRD_VARIANT_MF");
  RCHK(check_periodic(*prm, nb));
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return set_err(DRD_E_HIP, "no HIP device visible");
  if ((int64_t)(prm->nx + 3) * (prm->ny + 2) >= ((int64_t)1 << 31))
    return set_err(DRD_E_UNSUPPORTED, "grid too large for 32-bit indexing");
  drd_ctx* c = new drd_ctx();
  c->nb = nb;
  c->p = *prm;
  c->g = Grid::make(prm->nx, prm->ny, prm->lr_bd_type == 'p');
  c->s = make_scal(*prm);
  int rc = 0;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    rc = set_err(DRD_E_HIP, "hipStreamCreate failed");
  if (!rc) rc = ctx_alloc(c);
  if (!rc) rc = build_tables(c);
  if (!rc) rc = eager_vprec(c);
  if (rc) {
    ctx_free(c);
    delete c;
    return rc;
  }
  *out = c;
  return 0;
}

extern "C" int drd_init(const drd_params* prm, drd_ctx** out) { return ctx_create(prm, 1, out); }

extern "C" int drd_set_initial_guess(drd_ctx* c, int32_t mode) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  if (mode != 0 && mode != 1) return set_err(DRD_E_ARG, "initial-guess mode must be 0 or 1");
  c->guess = mode;
  return 0;
}

extern "C" int drd_mg_fused(drd_ctx* c, int32_t on) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);  // captured chunks hold the other form's launches
  c->mg_fused = on != 0;
  velo_mg_set_fused(c->vmg, c->mg_fused);
  velo_mg_set_fused(c->vmg1, c->mg_fused);
  return 0;
}

extern "C" int drd_use_graphs(drd_ctx* c, int32_t on) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!on) drop_graphs(c);
  c->use_graphs = on != 0;
  return 0;
}

extern "C" int drd_graph_count(drd_ctx* c, int32_t* live, int64_t* recaptured) {
  if (!c || !live || !recaptured) return set_err(DRD_E_ARG, "null argument");
  *live = (int32_t)c->graphs.size();
  *recaptured = c->graph_recaptures;
  return 0;
}

extern "C" int drd_set_params(drd_ctx* c, const drd_params* prm) {
  if (!c || !prm) return set_err(DRD_E_ARG, "null argument");
  if (prm->nx != c->p.nx || prm->ny != c->p.ny) return set_err(DRD_E_ARG, "grid size differs from the context");
  if (prm->lr_bd_type != 'n' && prm->lr_bd_type != 'd' && prm->lr_bd_type != 'p')
    return set_err(DRD_E_ARG, "lr_bd_type must be 'n', 'd' or 'p'");
  if ((prm->lr_bd_type == 'p') != (c->g.px != 0))
    return set_err(DRD_E_ARG, "periodic x is fixed when the context is created");
  if (prm->phi_mode != 1 && prm->phi_mode != 2) return set_err(DRD_E_ARG, "phi_mode must be 1 or 2");
  if (prm->variant != DRD_VARIANT_TP && prm->variant != DRD_VARIANT_MF)
    return set_err(DRD_E_ARG, "variant must be DRD_VARIANT_TP or DRD_VARIANT_MF");
  RCHK(check_periodic(*prm, c->nb));
  if (std::memcmp(prm, &c->p, sizeof(drd_params)) == 0) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->p = *prm;
  c->s = make_scal(*prm);
  hipFree(c->pre_lcc);
  hipFree(c->ph_lcc);
  c->pre_lcc = c->ph_lcc = nullptr;
  if (c->vprec && velo_prec_update(c->vprec, c->s))  // its eigenvalues depend on dt, re, dx, dz
    return set_err(DRD_E_HIP, "preconditioner tables could not be rebuilt");
  velo_mg_destroy(c->vmg);  // its levels' scalars follow the parameters; rebuilt when next used
  c->vmg = nullptr;
  velo_prec_destroy(c->vprec1);  // the BDF1 forms likewise
  c->vprec1 = nullptr;
  velo_mg_destroy(c->vmg1);
  c->vmg1 = nullptr;
  RCHK(build_tables(c));
  RCHK(eager_vprec(c));
  return 0;
}

extern "C" int drd_finalize(drd_ctx* c) {
  if (!c) return 0;
  hipStreamSynchronize(c->stream);
  ctx_free(c);
  delete c;
  return 0;
}

extern "C" int drd_synchronize(drd_ctx* c) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// Periodic x: a field of n doubles written from outside takes the periodic
// ghost columns (the kind follows from its size; 1-D fields are left alone).
static void fill_ghosts(drd_ctx* c, double* f, int64_t n) {
  if (!c->g.px) return;
  const Grid& g = c->g;
  const int kind = n == g.nuf() ? 1 : (n == g.nvf() ? 2 : (n == g.nc() ? 0 : -1));
  if (kind >= 0) launch_xghost_fill(c->stream, g, f, kind);
}

extern "C" int drd_set_field(drd_ctx* c, int32_t field, const double* host) {
  if (!c || !host) return set_err(DRD_E_ARG, "null argument");
  const int64_t n = drd_field_size(&c->p, field);
  if (n <= 0) return set_err(DRD_E_ARG, "unknown field id");
  HIPCHK(hipMemcpyAsync(c->f[field], host, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  fill_ghosts(c, c->f[field], n);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (field == DRD_F_ANGLEB || field == DRD_F_ANGLET) RCHK(upload_angles(c));
  return 0;
}

extern "C" int drd_get_field(drd_ctx* c, int32_t field, double* host) {
  if (!c || !host) return set_err(DRD_E_ARG, "null argument");
  const int64_t n = drd_field_size(&c->p, field);
  if (n <= 0) return set_err(DRD_E_ARG, "unknown field id");
  HIPCHK(hipMemcpyAsync(host, c->f[field], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static Coef coef_of(const drd_ctx* c) {
  Coef k{};
  k.rho = c->f[DRD_F_RHO];
  k.eta = c->f[DRD_F_ETA];
  // the BDF2 penalty mask: bdr_ph_penalty (TP), the smooth bdr_ph (MF/VeloSolve_2_combined.f90:580, :613)
  k.chi = c->p.variant == DRD_VARIANT_MF ? c->f[DRD_F_BDR_PH] : c->f[DRD_F_BDR_PH_PENALTY];
  k.bdr = c->f[DRD_F_BDR_PH];
  k.bu = c->f[DRD_F_BDR_PH_USED];
  k.bg = c->f[DRD_F_BDR_PH_GRAD];
  k.minner = c->f[DRD_F_PH_M_INNER];
  k.m0 = c->f[DRD_F_PH_M_0];
  return k;
}

static int64_t op_size(const drd_ctx* c, int op) {
  switch (op) {
    case DRD_OP_VELO1:
    case DRD_OP_VELO2:
      return c->g.nuv();
    case DRD_OP_PHMU:
      return 2 * c->g.ncell();
    case DRD_OP_PH:
    case DRD_OP_PHORI:
      return c->g.ncell();
    case kOpExt:
      return c->ext_n;
  }
  return -1;
}

// Segments of the operator's Krylov vectors: the phi|mu vector is summed as
// two blocks (its kernels produce the phi and mu rows of a cell together).
static int op_nseg(int op) { return op == DRD_OP_PHMU ? 2 : 1; }

// resources of PhSolve, its VCOST preconditioner and ESOLP, made on first use
static int aux_ready(drd_ctx* c) {
  if (c->aux_work) return 0;
  const Grid& g = c->g;
  c->d1_nx = dct1_table_create(g.nx);
  c->d1_ny = dct1_table_create(g.ny);
  c->d1_nx1 = dct1_table_create(g.nx + 1);
  if (!c->d1_nx || !c->d1_ny || !c->d1_nx1) return set_err(DRD_E_HIP, "VCOST tables");
  RCHK(dalloc(&c->ph_f2, g.nc()));
  RCHK(dalloc(&c->ph_t, g.ncell()));
  RCHK(dalloc(&c->aux_work, std::max<int64_t>(2 * g.ncell(), 3 * g.nuv())));
  return 0;
}

static int read_state(drd_ctx* c);

// drd_bicgstab_cb's operator: the caller's MATVEC(ALPHA, X, BETA, Y) on host
// copies, called as LIB/BiCGSTABREVCOM.f calls it (r0 = -1 A x0 + 1 b; the
// iteration's products with beta 0 into a zeroed vector), then the fused
// modes' reductions on the device.  A fused mode first reads the solve's
// state (every queued kernel has run by then), and a finished solve or an
// early exit calls no MATVEC, as the reference's loop would not.
static int ext_host(drd_ctx* c) {
  if (c->ext_h0) return 0;
  HIPCHK(hipHostMalloc((void**)&c->ext_h0, (size_t)c->nmax * sizeof(double), hipHostMallocDefault));
  HIPCHK(hipHostMalloc((void**)&c->ext_h1, (size_t)c->nmax * sizeof(double), hipHostMallocDefault));
  return 0;
}
static int ext_op(drd_ctx* c, int mode, const double* x, double* y, double alpha, double beta, double* a0,
                  const double* b, KState* ks) {
  if (!c->ext_mv) return set_err(DRD_E_ARG, "no MATVEC outside drd_bicgstab_cb");
  const int64_t n = c->ext_n;
  const size_t bytes = (size_t)n * sizeof(double);
  RCHK(ext_host(c));
  double *hx = c->ext_h0, *hy = c->ext_h1;
  bool call = true;
  if (mode != MV_PLAIN) {
    RCHK(read_state(c));
    const KState& st = *c->ks_host;
    call = !st.done && !st.early;
    if (call && mode == MV_INIT && !st.xnz) {  // x0 = 0: r0 = b (REVCOM:229-240)
      launch_copy(c->stream, n, b, y);
      call = false;
    }
  }
  if (call) {
    HIPCHK(hipMemcpyAsync(hx, x, bytes, hipMemcpyDeviceToHost, c->stream));
    const double* yin = mode == MV_INIT ? b : (mode == MV_PLAIN && beta != 0.0 ? y : nullptr);
    if (yin) HIPCHK(hipMemcpyAsync(hy, yin, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!yin) std::memset(hy, 0, bytes);
    c->ext_mv(&alpha, hx, &beta, hy);
    HIPCHK(hipMemcpyAsync(y, hy, bytes, hipMemcpyHostToDevice, c->stream));
  }
  if (mode != MV_PLAIN) launch_ext_mv(c->stream, mode, n, y, a0, c->part, ks);
  return 0;
}
// drd_bicgstab_cb's preconditioner: the caller's PSOLVE(X, B) on host copies
static int ext_prec(drd_ctx* c, const double* b, double* x) {
  const size_t bytes = (size_t)c->ext_n * sizeof(double);
  RCHK(ext_host(c));
  RCHK(read_state(c));
  const KState& st = *c->ks_host;
  if (st.done || st.early) return 0;  // the queued kernels skip as well
  HIPCHK(hipMemcpyAsync(c->ext_h1, b, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  c->ext_ps(c->ext_h0, c->ext_h1);
  HIPCHK(hipMemcpyAsync(x, c->ext_h0, bytes, hipMemcpyHostToDevice, c->stream));
  return 0;
}

// y = alpha*A*x + beta*y (device pointers), or a fused Krylov mode
static int apply_op(drd_ctx* c, int op, int mode, const double* x, double* y, double alpha, double beta,
                    double* a0, const double* b, KState* ks) {
  const Coef k = coef_of(c);
  switch (op) {
    case DRD_OP_PH:
      RCHK(aux_ready(c));
      launch_ph_matvec(c->stream, mode, c->g, c->s, k.m0, k.minner, x, c->ph_t, y, alpha, beta, a0, b, c->part,
                       ks);
      break;
    case DRD_OP_VELO1:
    case DRD_OP_VELO2: {
      const bool timed = op == DRD_OP_VELO2 && mode == MV_DOT1 && c->tm_which == DRD_K_VELO_MATVEC &&
                         2 * (size_t)c->tm_n < c->tm_ev.size();
      launch_velo_matvec(c->stream, op == DRD_OP_VELO1 ? 1 : 2, mode, c->g, c->s, k, x, y, alpha, beta, a0, b,
                         c->part, ks, true, timed ? c->tm_ev[2 * c->tm_n] : nullptr,
                         timed ? c->tm_ev[2 * c->tm_n + 1] : nullptr);
      if (timed) ++c->tm_n;
      break;
    }
    case DRD_OP_PHMU:
      launch_phmu_matvec(c->stream, mode, c->g, c->s, k, x, y, alpha, beta, a0, b, c->part, ks);
      break;
    case kOpExt:
      RCHK(ext_op(c, mode, x, y, alpha, beta, a0, b, ks));
      break;
    case DRD_OP_PHORI:
      if (c->nb != 1) return set_err(DRD_E_UNSUPPORTED, "PhSolve_ori takes a single-instance context");
      RCHK(aux_ready(c));
      launch_phori_matvec(c->stream, mode, c->g, c->s, c->phori_cb, c->phori_ct, x, c->ph_t, y, alpha, beta, a0, b,
                          c->part, ks);
      break;
    default:
      return set_err(DRD_E_UNSUPPORTED, "operator not built");
  }
  return 0;
}

// Prepares a preconditioner once per solve (Jacobi diagonal, DCT scaling).
static PhmuCoef32 phmu_c32(const drd_ctx* c) {
  const int64_t n = c->g.nc();
  const float* f = c->rb_c32;
  return PhmuCoef32{f, f + n, f + 2 * n, f + 3 * n, f + 4 * n, f + 5 * n};
}

static int prec_setup(drd_ctx* c, int prec) {
  if (prec == kPsExt) return 0;
  if (c->g.px && !periodic_prec_ok(prec))
    return set_err(DRD_E_UNSUPPORTED, "periodic x takes the Jacobi, identity and phi/mu block preconditioners");
  if (prec == DRD_PS_VELO1_JACOBI) launch_velo_jacobi_diag(c->stream, c->g, c->s, coef_of(c), c->diag, c->s.re);
  if (prec == DRD_PS_VELO2_JACOBI)
    launch_velo_jacobi_diag(c->stream, c->g, c->s, coef_of(c), c->diag, 1.5 * c->s.re);
  if (prec == DRD_PS_PHMU_RBGS32) launch_phmu_coef32(c->stream, c->g, c->s, coef_of(c), phmu_c32(c));
  if (prec == DRD_PS_VELO_DCT32 || prec == DRD_PS_VELO_DCT64 || prec == DRD_PS_VELO_LEARNED) {
    const bool fp32 = (prec != DRD_PS_VELO_DCT64);
    if (c->vprec && c->vprec_fp32 != fp32) {
      HIPCHK(hipStreamSynchronize(c->stream));
      drop_graphs(c);
      velo_prec_destroy(c->vprec);
      c->vprec = nullptr;
    }
    if (!c->vprec) {
      int err = 0;
      c->vprec = make_vprec(c, fp32, &err);
      if (!c->vprec)
        return set_err(err ? err : DRD_E_HIP, "DCT velocity preconditioner needs power-of-two nx, ny <= 4096");
      c->vprec_fp32 = fp32;
    }
    velo_prec_setup(c->stream, c->vprec, c->g, c->s, coef_of(c));
  }
  if (prec == DRD_PS_VELO_MG || prec == DRD_PS_VELO_MG32) {
    const bool f32 = prec == DRD_PS_VELO_MG32;
    if (c->vmg && velo_mg_f32(c->vmg) != f32) {  // captured chunks point into the old hierarchy
      HIPCHK(hipStreamSynchronize(c->stream));
      drop_graphs(c);
      velo_mg_destroy(c->vmg);
      c->vmg = nullptr;
    }
    if (!c->vmg) {
      int err = 0;
      c->vmg = velo_mg_create(c->g, c->s, f32, &err, c->nb > 1 ? c->mg_mem : nullptr);
      if (!c->vmg)
        return set_err(err ? err : DRD_E_HIP,
                       "multigrid velocity preconditioner needs even nx >= 128, ny >= 64 (and a coarsest grid of at most 256 x 256 cells or of powers of two)");
    }
    velo_mg_set_fused(c->vmg, c->mg_fused);
    velo_mg_setup(c->stream, c->vmg, coef_of(c));
  }
  if (prec == DRD_PS_VELO1_DCT32 || prec == DRD_PS_VELO1_DCT64) {
    const bool fp32 = prec == DRD_PS_VELO1_DCT32;
    const Scal s1 = bdf1_scal(c->s);
    if (c->vprec1 && c->vprec1_fp32 != fp32) {
      HIPCHK(hipStreamSynchronize(c->stream));
      drop_graphs(c);
      velo_prec_destroy(c->vprec1);
      c->vprec1 = nullptr;
    }
    if (!c->vprec1) {
      int err = 0;
      const VeloPrecBufs bufs{c->vp_inv64, c->vp_inv32, c->vp_work};
      c->vprec1 = velo_prec_create(c->g, s1, fp32, &err, &bufs);
      if (!c->vprec1)
        return set_err(err ? err : DRD_E_HIP, "DCT velocity preconditioner needs power-of-two nx, ny <= 4096");
      c->vprec1_fp32 = fp32;
    }
    velo_prec_setup(c->stream, c->vprec1, c->g, s1, coef_of(c));
  }
  if (prec == DRD_PS_VELO1_MG || prec == DRD_PS_VELO1_MG32) {
    const bool f32 = prec == DRD_PS_VELO1_MG32;
    if (c->vmg1 && velo_mg_f32(c->vmg1) != f32) {
      HIPCHK(hipStreamSynchronize(c->stream));
      drop_graphs(c);
      velo_mg_destroy(c->vmg1);
      c->vmg1 = nullptr;
    }
    if (!c->vmg1) {
      int err = 0;
      c->vmg1 = velo_mg_create(c->g, bdf1_scal(c->s), f32, &err, c->nb > 1 ? c->mg1_mem : nullptr);
      if (!c->vmg1)
        return set_err(err ? err : DRD_E_HIP,
                       "multigrid velocity preconditioner needs even nx >= 128, ny >= 64 (and a coarsest grid of at most 256 x 256 cells or of powers of two)");
    }
    velo_mg_set_fused(c->vmg1, c->mg_fused);
    velo_mg_setup(c->stream, c->vmg1, coef_of(c));
  }
  return 0;
}

// z += alpha s_d N: the learned correction for the vector b (z = M b on
// entry), alpha minimising |b - A z| (device-side scalars, no host sync), so
// the corrected residual is never larger than the plain DCT solve's.
static int learn_correct(drd_ctx* c, const double* b, double* z) {
  Learn* L = c->learn;
  hipStream_t st = c->stream;
  const int64_t n = c->g.nuv();
  double* d = L->dac;       // d = b - A z
  double* ac = L->dac + n;  // A (correction)
  launch_copy(st, n, b, d);
  RCHK(apply_op(c, DRD_OP_VELO2, MV_PLAIN, z, d, -1.0, 1.0, nullptr, nullptr, nullptr));
  launch_norms3(st, n, b, d, z, L->ss_part, L->ss);
  launch_pack_features(st, c->g, b, d, z, c->f[DRD_F_ETA], c->f[DRD_F_BDR_PH_PENALTY], L->ss, L->img);
  RCHK(unet_forward(st, L->net, L->img, L->out));
  launch_unpack_correction(st, c->g, L->out, L->ss, L->w);
  RCHK(apply_op(c, DRD_OP_VELO2, MV_PLAIN, L->w, ac, 1.0, 0.0, nullptr, nullptr, nullptr));
  launch_multidot(st, n, 1, ac, L->dac, n, 2, L->part, L->dots);  // (Ac, d), (Ac, Ac)
  launch_axpy_ratio(st, n, L->dots, L->w, z);
  return 0;
}

// x_out = M^-1 b_in; may return b_in itself (identity).
static int apply_prec(drd_ctx* c, int prec, int op, const double* b, double* x, const double** res,
                      const KState* ks) {
  switch (prec) {
    case kPsExt:
      if (op != kOpExt || !c->ext_ps) break;
      RCHK(ext_prec(c, b, x));
      *res = x;
      return 0;
    case DRD_PS_VELO_LEARNED:
      if (op != DRD_OP_VELO2 || !c->vprec) break;
      if (!c->learn || c->learn->train_iters == 0)
        return set_err(DRD_E_STATE, "learned preconditioner: network not trained (drd_learn_train)");
      velo_prec_apply(c->stream, c->vprec, c->g, b, x, ks);
      RCHK(learn_correct(c, b, x));
      *res = x;
      return 0;
    case DRD_PS_IDENTITY:
      *res = b;
      return 0;
    case DRD_PS_VELO1_JACOBI:
    case DRD_PS_VELO2_JACOBI:
      launch_div_diag(c->stream, op_size(c, op), b, c->diag, x, ks);
      *res = x;
      return 0;
    case DRD_PS_PHMU_JACOBI:
      launch_phmu_jacobi(c->stream, c->g, c->s, b, x, ks);
      *res = x;
      return 0;
    case DRD_PS_PHMU_RBGS:
      if (op != DRD_OP_PHMU) break;
      launch_phmu_rbgs(c->stream, c->g, c->s, coef_of(c), nullptr, b, x, c->rb_work, ks);
      *res = x;
      return 0;
    case DRD_PS_PHMU_RBGS32: {  // the FP32 coefficients are those of its last prec_setup
      if (op != DRD_OP_PHMU) break;
      const PhmuCoef32 c32 = phmu_c32(c);
      const KTimer tm{c->tm_ev.data(), (int)(c->tm_ev.size() / 2), &c->tm_n};
      launch_phmu_rbgs(c->stream, c->g, c->s, coef_of(c), &c32, b, x, c->rb_work, ks,
                       c->tm_which == DRD_K_PHMU_SWEEP32 ? &tm : nullptr);
      *res = x;
      return 0;
    }
    case DRD_PS_PHMU_BLOCK:
      if (op != DRD_OP_PHMU) break;
      launch_phmu_block(c->stream, c->g, c->s, coef_of(c), b, x, ks);
      *res = x;
      return 0;
    case DRD_PS_VELO_DCT32:
    case DRD_PS_VELO_DCT64:
      if (op != DRD_OP_VELO2 || !c->vprec) break;
      velo_prec_apply(c->stream, c->vprec, c->g, b, x, ks);
      *res = x;
      return 0;
    case DRD_PS_VELO_MG:
    case DRD_PS_VELO_MG32:
      if (op != DRD_OP_VELO2 || !c->vmg || velo_mg_f32(c->vmg) != (prec == DRD_PS_VELO_MG32)) break;
      velo_mg_apply(c->stream, c->vmg, b, x, ks);
      *res = x;
      return 0;
    case DRD_PS_VELO1_DCT32:
    case DRD_PS_VELO1_DCT64:
      if (op != DRD_OP_VELO1 || !c->vprec1 || c->vprec1_fp32 != (prec == DRD_PS_VELO1_DCT32)) break;
      velo_prec_apply(c->stream, c->vprec1, c->g, b, x, ks);
      *res = x;
      return 0;
    case DRD_PS_VELO1_MG:
    case DRD_PS_VELO1_MG32:
      if (op != DRD_OP_VELO1 || !c->vmg1 || velo_mg_f32(c->vmg1) != (prec == DRD_PS_VELO1_MG32)) break;
      velo_mg_apply(c->stream, c->vmg1, b, x, ks);
      *res = x;
      return 0;
    case DRD_PS_PH_DCT:  // PSolve_CG MF/PhSolve.f90:296-313 = PSolve_CG_ori TP/PhSolve_ori.f90:402-416
      if (op != DRD_OP_PH && op != DRD_OP_PHORI) break;
      RCHK(aux_ready(c));
      launch_ph_dct_prec(c->stream, c->g, c->d1_nx, c->d1_ny, c->ph_lcc, b, x, c->aux_work, ks);
      *res = x;
      return 0;
    case DRD_PS_VELO_ESOLP:
      if (op != DRD_OP_VELO1 && op != DRD_OP_VELO2) break;
      RCHK(aux_ready(c));
      launch_esolp(c->stream, c->g, c->d1_nx1, c->d1_nx, c->esolp_tabs, b, x, c->aux_work, ks);
      *res = x;
      return 0;
  }
  return set_err(DRD_E_UNSUPPORTED, "preconditioner not built for this operator");
}

// every instance's KState: device <-> pinned host (one strided copy)
static int read_state(drd_ctx* c) {
  if (c->nb == 1)
    HIPCHK(hipMemcpyAsync(c->ks_host, c->ks, sizeof(KState), hipMemcpyDeviceToHost, c->stream));
  else
    HIPCHK(hipMemcpy2DAsync(c->ks_host, sizeof(KState), c->ks, (size_t)c->bs, sizeof(KState), (size_t)c->nb,
                            hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}
static int put_state(drd_ctx* c, const KState& h) {
  for (int i = 0; i < c->nb; ++i) c->ks_host[i] = h;
  if (c->nb == 1)
    HIPCHK(hipMemcpyAsync(c->ks, c->ks_host, sizeof(KState), hipMemcpyHostToDevice, c->stream));
  else
    HIPCHK(hipMemcpy2DAsync(c->ks, (size_t)c->bs, c->ks_host, sizeof(KState), sizeof(KState), (size_t)c->nb,
                            hipMemcpyHostToDevice, c->stream));
  return 0;
}
static bool all_done(const drd_ctx* c) {
  for (int i = 0; i < c->nb; ++i)
    if (!c->ks_host[i].done) return false;
  return true;
}

// FP64 ||b - A x|| / ||b|| (true residual; an addition to the reference)
static int true_resid(drd_ctx* c, int op, int64_t n, const double* b, const double* x, double* out) {
  launch_copy(c->stream, n, b, c->tmp2);
  RCHK(apply_op(c, op, MV_PLAIN, x, c->tmp2, 1.0, 0.0, nullptr, nullptr, nullptr));
  launch_resid_norms(c->stream, n, op_nseg(op), b, c->tmp2, c->part, c->ks);
  RCHK(read_state(c));
  const double bn = std::sqrt(c->ks_host->red[1]);
  *out = std::sqrt(c->ks_host->red[0]) / (bn == 0.0 ? 1.0 : bn);
  return 0;
}

// Everything a captured BiCGSTAB chunk bakes into its kernels besides the
// GraphKey: FNV-1a over the workspace, coefficient and preconditioner
// pointers, the scalars, the grid and the batch group.
static uint64_t graph_print(const drd_ctx* c, int prec) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](const void* p, size_t n) {
    const unsigned char* q = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) h = (h ^ q[i]) * 1099511628211ull;
  };
  auto ptr = [&](const void* p) { mix(&p, sizeof(p)); };
  for (int f = 0; f < DRD_F_COUNT; ++f)
    if (f != DRD_F_U && f != DRD_F_V && f != DRD_F_U_PRE && f != DRD_F_V_PRE && f != DRD_F_PH &&
        f != DRD_F_PH_PRE && f != DRD_F_RMU && f != DRD_F_RMU_PRE && f != DRD_F_PSI1 && f != DRD_F_PSI2)
      ptr(c->f[f]);  // the coefficient fields (the rotated state fields are not read by a chunk)
  for (double* w : c->kw) ptr(w);
  ptr(c->part);
  ptr(c->ks);
  ptr(c->diag);
  ptr(c->rb_work);
  ptr(c->rb_c32);
  // the preconditioner object the chunk's kernels read (others may come and go)
  if (prec == DRD_PS_VELO_DCT32 || prec == DRD_PS_VELO_DCT64) ptr(c->vprec);
  if (prec == DRD_PS_VELO_MG || prec == DRD_PS_VELO_MG32) ptr(c->vmg);
  if (prec == DRD_PS_VELO1_DCT32 || prec == DRD_PS_VELO1_DCT64) ptr(c->vprec1);
  if (prec == DRD_PS_VELO1_MG || prec == DRD_PS_VELO1_MG32) ptr(c->vmg1);
  ptr(c->vp_work);
  mix(&c->vprec_fp32, sizeof(c->vprec_fp32));
  mix(&c->vprec1_fp32, sizeof(c->vprec1_fp32));
  mix(&c->s, sizeof(c->s));
  mix(&c->g, sizeof(c->g));
  const BatchCfg bc = batch_cfg();
  mix(&bc.n, sizeof(bc.n));
  mix(&bc.bs, sizeof(bc.bs));
  return h;
}

// Ends a stream capture left open by an error return, so the stream is
// usable again and no half-captured graph survives.
struct CaptureGuard {
  hipStream_t st;
  bool active = false;
  ~CaptureGuard() {
    if (!active) return;
    hipGraph_t gr = nullptr;
    if (hipStreamEndCapture(st, &gr) == hipSuccess && gr) (void)hipGraphDestroy(gr);
    (void)hipGetLastError();
  }
};

// BICGSTAB LIB/BiCGSTAB.f + LIB/BiCGSTABREVCOM.f on device vectors b, x.
// Optional recording: x after iteration snap_at[j] is copied to snaps + j*n
// (nsnap entries, increasing).
static int run_bicgstab(drd_ctx* c, int op, int prec, int64_t n, const double* b, double* x, int maxit,
                        double tol, drd_solve_info* info, bool want_true, const int* snap_at = nullptr,
                        int nsnap = 0, double* snaps = nullptr) {
  if (n != op_size(c, op)) return set_err(DRD_E_ARG, "vector length does not match the operator");
  if (prec == DRD_PS_VELO_LEARNED)
    return set_err(DRD_E_UNSUPPORTED, "the learned preconditioner is not linear: use FGMRES");
  if (maxit <= 0) {  // LIB/BiCGSTAB.f:161-163
    if (info) *info = drd_solve_info{maxit, -3, tol, 0.0};
    return 0;
  }
  double *R = c->kw[0], *RT = c->kw[1], *P = c->kw[2], *V = c->kw[3], *T = c->kw[4], *PHb = c->kw[5],
         *SHb = c->kw[6];
  KState h{};
  h.tol = tol;
  h.maxit = maxit;
  h.iter = maxit;
  h.resid = tol;
  h.rho1 = 1.0;
  h.omega = 1.0;
  const double eps = std::ldexp(1.0, -53);  // DLAMCH('E')
  h.breaktol = eps * eps;                   // LIB/getbreak.f:11-12
  RCHK(put_state(c, h));
  RCHK(prec_setup(c, prec));
  const int ns = op_nseg(op);
  const bool fused_block = (op == DRD_OP_PHMU && prec == DRD_PS_PHMU_BLOCK);
  launch_init_norms(c->stream, n, ns, x, b, c->part, c->ks);
  RCHK(apply_op(c, op, MV_INIT, x, R, -1.0, 1.0, RT, b, c->ks));
  // chunks of iterations are queued without host round trips; with graphs
  // on, a chunk of length >= 2 is one replay of a captured hipGraph
  const bool graphable = c->use_graphs && nsnap == 0 && c->tm_which < 0 && op != kOpExt &&
                         (prec == DRD_PS_IDENTITY || prec == DRD_PS_VELO1_JACOBI || prec == DRD_PS_VELO2_JACOBI ||
                          prec == DRD_PS_PHMU_JACOBI ||
                          prec == DRD_PS_PHMU_BLOCK || prec == DRD_PS_VELO_DCT32 || prec == DRD_PS_VELO_DCT64 ||
                          prec == DRD_PS_VELO_MG || prec == DRD_PS_VELO_MG32 || prec == DRD_PS_PHMU_RBGS ||
                          prec == DRD_PS_PHMU_RBGS32 || prec == DRD_PS_VELO1_DCT32 ||
                          prec == DRD_PS_VELO1_DCT64 || prec == DRD_PS_VELO1_MG || prec == DRD_PS_VELO1_MG32);
  // The first chunk covers one iteration less than the previous solve of
  // this operator took (solves of successive steps need similar counts);
  // later chunks grow 1, 2, 4, ... 32.  A chunk queued past convergence
  // costs one dispatch per kernel, so the tail a power-of-two schedule
  // leaves (up to half the solve) is what this saves.
  const int pred = (op >= 0 && op < 5) ? c->pred_its[op] : 0;
  int chunk = pred > 2 ? std::min(pred - 1, maxit) : 1, after = 1, total = 0, next_snap = 0;
  const uint64_t print = graphable ? graph_print(c, prec) : 0;
  for (;;) {
    hipGraphExec_t gx = nullptr;
    CaptureGuard cap{c->stream};
    if (graphable && chunk >= 2) {
      const GraphKey key{op, prec, chunk, b, x};
      auto hit = c->graphs.find(key);
      if (hit != c->graphs.end() && hit->second.print != print) {  // stale: capture it again
        (void)hipGraphExecDestroy(hit->second.exec);
        c->graphs.erase(hit);
        hit = c->graphs.end();
        ++c->graph_recaptures;
      }
      if (hit != c->graphs.end()) {
        gx = hit->second.exec;
      } else {
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        cap.active = true;
      }
    }
    for (int it = 0; it < (gx ? 0 : chunk); ++it) {
      const double* ph = nullptr;
      const double* sh = nullptr;
      if (fused_block) {  // p, phat = B^-1 p in one sweep
        launch_bicg_p_block(c->stream, c->g, c->s, coef_of(c), R, V, P, PHb, c->ks);
        ph = PHb;
      } else {
        launch_bicg_p(c->stream, n, R, V, P, c->ks);
        RCHK(apply_prec(c, prec, op, P, PHb, &ph, c->ks));
      }
      RCHK(apply_op(c, op, MV_DOT1, ph, V, 1.0, 0.0, RT, nullptr, c->ks));
      // after an early exit (||s|| <= tol) the kernels up to the x, r sweep
      // skip, and that sweep makes the exit's update of x
      if (fused_block) {  // s, ||s||^2, shat = B^-1 s in one sweep
        launch_bicg_s_block(c->stream, c->g, c->s, coef_of(c), R, V, SHb, c->part, c->ks);
        sh = SHb;
      } else {
        launch_bicg_s(c->stream, n, ns, R, V, c->part, c->ks);
        RCHK(apply_prec(c, prec, op, R, SHb, &sh, c->ks));
      }
      RCHK(apply_op(c, op, MV_DOT2, sh, T, 1.0, 0.0, R, nullptr, c->ks));
      launch_bicg_xr(c->stream, n, ns, x, R, RT, ph, sh, T, c->part, c->ks);
      if (next_snap < nsnap && total + it + 1 == snap_at[next_snap]) {
        launch_copy(c->stream, n, x, snaps + (size_t)next_snap * n);
        ++next_snap;
      }
    }
    if (cap.active) {
      hipGraph_t gr = nullptr;
      cap.active = false;
      HIPCHK(hipStreamEndCapture(c->stream, &gr));
      const hipError_t e = hipGraphInstantiate(&gx, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      HIPCHK(e);
      c->graphs[GraphKey{op, prec, chunk, b, x}] = GraphEntry{gx, print};
    }
    if (gx) HIPCHK(hipGraphLaunch(gx, c->stream));
    HIPCHK(hipGetLastError());
    total += chunk;
    RCHK(read_state(c));
    if (all_done(c) || total > maxit + 2) break;
    chunk = after;
    after = std::min(after * 2, 32);
  }
  if (!all_done(c)) return set_err(DRD_E_STATE, "BiCGSTAB state machine did not terminate");
  int most = 0;
  for (int i = 0; i < c->nb; ++i) {
    c->last[i] = drd_solve_info{c->ks_host[i].iter, c->ks_host[i].info, c->ks_host[i].resid, 0.0};
    most = std::max(most, (int)c->ks_host[i].iter);
  }
  if (op >= 0 && op < 5) c->pred_its[op] = most;
  const KState& s = *c->ks_host;
  if (info) {
    info->iter = s.iter;
    info->info = s.info;
    info->resid = s.resid;
    info->true_resid = 0.0;
    if (want_true) RCHK(true_resid(c, op, n, b, x, &info->true_resid));
  }
  return 0;
}

// ---- FGMRES(m) ----------------------------------------------------------------
constexpr int kFgmresRestart = 30;  // restart length used by the velocity solve

// Restarted flexible GMRES with right preconditioning (Saad, SIAM J. Sci.
// Comput. 14(2), 1993).  Orthogonalisation by classical Gram-Schmidt with a
// second pass when the first one cancels more than half of the norm
// ("twice is enough"); the Hessenberg matrix, Givens rotations and the
// small triangular solve live on the host, one synchronisation per inner
// iteration.  Stops when the Arnoldi residual estimate reaches tol*||b||,
// then confirms with the true residual (restarting if needed).
static int fgmres_alloc(drd_ctx* c, int m) {
  if (c->fg_m >= m && c->fg_V) return 0;
  hipFree(c->fg_V);
  hipFree(c->fg_Z);
  hipFree(c->fg_h);
  hipFree(c->fg_part);
  if (c->fg_hh) hipHostFree(c->fg_hh);
  c->fg_V = c->fg_Z = c->fg_h = c->fg_part = nullptr;
  c->fg_hh = nullptr;
  c->fg_m = 0;
  RCHK(dalloc(&c->fg_V, (int64_t)(m + 1) * c->nmax));
  RCHK(dalloc(&c->fg_Z, (int64_t)m * c->nmax));
  RCHK(dalloc(&c->fg_h, 4 * (int64_t)(m + 2)));
  RCHK(dalloc(&c->fg_part, 2 * (int64_t)(m + 2) * c->tiles));
  HIPCHK(hipHostMalloc((void**)&c->fg_hh, 4 * (m + 2) * sizeof(double), hipHostMallocDefault));
  c->fg_m = m;
  return 0;
}

static int run_fgmres(drd_ctx* c, int op, int prec, int64_t n, const double* b, double* x, int m, int maxit,
                      double tol, drd_solve_info* info) {
  if (n != op_size(c, op)) return set_err(DRD_E_ARG, "vector length does not match the operator");
  if (m < 1 || m > 200) return set_err(DRD_E_ARG, "restart must be in [1, 200]");
  RCHK(fgmres_alloc(c, m));
  RCHK(prec_setup(c, prec));
  double* V = c->fg_V;
  double* Z = c->fg_Z;
  const int64_t ld = n;
  double* hd = c->fg_h;  // [0, m+2): first pass, [m+2, 2m+4): second pass, then norms
  double* hh = c->fg_hh;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m);
  auto sync_read = [&](int count) -> int {
    HIPCHK(hipMemcpyAsync(hh, hd, count * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  };
  // ||b||
  const int ns = op_nseg(op);
  double* fpart = c->fg_part;
  launch_multidot(c->stream, n, ns, b, b, ld, 1, fpart, hd);
  RCHK(sync_read(1));
  const double bn = std::sqrt(hh[0]);
  drd_solve_info loc{0, 1, tol, 0.0};
  if (bn == 0.0) {
    HIPCHK(hipMemsetAsync(x, 0, n * sizeof(double), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    loc = drd_solve_info{0, 0, 0.0, 0.0};
    if (info) *info = loc;
    return 0;
  }
  int it = 0;
  for (;;) {
    // r = b - A x into V_0, beta = ||r||
    launch_copy(c->stream, n, b, V);
    RCHK(apply_op(c, op, MV_PLAIN, x, V, -1.0, 1.0, nullptr, nullptr, nullptr));
    launch_multidot(c->stream, n, ns, V, V, ld, 1, fpart, hd);
    RCHK(sync_read(1));
    const double beta = std::sqrt(hh[0]);
    loc.resid = beta / bn;
    loc.true_resid = loc.resid;
    if (loc.resid <= tol) {
      loc.info = 0;
      break;
    }
    if (it >= maxit) {
      loc.info = 1;
      break;
    }
    launch_scale_into(c->stream, n, V, beta, V);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int jdone = 0;
    for (int j = 0; j < m && it < maxit; ++j) {
      double* vj = V + (size_t)j * ld;
      double* zj = Z + (size_t)j * ld;
      double* w = V + (size_t)(j + 1) * ld;
      const double* zr = nullptr;
      // the learned correction opens each cycle; the cycle's other directions
      // come from the plain FP32 DCT solve
      const int ev = c->learn ? c->learn->every : 0;
      const bool corr = ev > 0 ? j % ev == 0 : j == 0;
      const int pj = (prec == DRD_PS_VELO_LEARNED && !corr) ? DRD_PS_VELO_DCT32 : prec;
      RCHK(apply_prec(c, pj, op, vj, zj, &zr, nullptr));
      if (zr != zj) launch_copy(c->stream, n, zr, zj);
      RCHK(apply_op(c, op, MV_PLAIN, zj, w, 1.0, 0.0, nullptr, nullptr, nullptr));
      // pass 1: h = V^T w and ||w||^2 (w = V_{j+1} is contiguous with V)
      launch_multidot(c->stream, n, ns, w, V, ld, j + 2, fpart, hd);
      launch_multiaxpy(c->stream, n, ns, w, V, ld, j + 1, hd, -1.0, fpart, hd + 2 * (m + 2), 1);
      RCHK(sync_read(2 * (m + 2) + 1));
      double* hcol = &H[(size_t)(m + 1) * j];
      for (int i = 0; i <= j; ++i) hcol[i] = hh[i];
      const double before = hh[j + 1];
      double after = hh[2 * (m + 2)];
      if (after < 0.25 * before) {  // pass 2
        launch_multidot(c->stream, n, ns, w, V, ld, j + 1, fpart, hd + (m + 2));
        launch_multiaxpy(c->stream, n, ns, w, V, ld, j + 1, hd + (m + 2), -1.0, fpart, hd + 2 * (m + 2),
                         1);
        RCHK(sync_read(2 * (m + 2) + 1));
        for (int i = 0; i <= j; ++i) hcol[i] += hh[(m + 2) + i];
        after = hh[2 * (m + 2)];
      }
      const double hn = std::sqrt(after);
      hcol[j + 1] = hn;
      if (hn != 0.0) launch_scale_into(c->stream, n, w, hn, w);
      for (int i = 0; i < j; ++i) {
        const double t1 = cs[i] * hcol[i] + sn[i] * hcol[i + 1];
        hcol[i + 1] = -sn[i] * hcol[i] + cs[i] * hcol[i + 1];
        hcol[i] = t1;
      }
      const double den = std::hypot(hcol[j], hcol[j + 1]);
      cs[j] = den == 0.0 ? 1.0 : hcol[j] / den;
      sn[j] = den == 0.0 ? 0.0 : hcol[j + 1] / den;
      hcol[j] = den;
      hcol[j + 1] = 0.0;
      g[j + 1] = -sn[j] * g[j];
      g[j] = cs[j] * g[j];
      ++it;
      jdone = j + 1;
      if (std::fabs(g[j + 1]) <= tol * bn || hn == 0.0) break;
    }
    for (int i = jdone - 1; i >= 0; --i) {
      double s = g[i];
      for (int k = i + 1; k < jdone; ++k) s -= H[i + (size_t)(m + 1) * k] * y[k];
      y[i] = s / H[i + (size_t)(m + 1) * i];
    }
    for (int i = 0; i < jdone; ++i) hh[i] = y[i];
    HIPCHK(hipMemcpyAsync(hd, hh, jdone * sizeof(double), hipMemcpyHostToDevice, c->stream));
    launch_multiaxpy(c->stream, n, ns, x, Z, ld, jdone, hd, 1.0, fpart, nullptr, 0);
  }
  HIPCHK(hipGetLastError());
  loc.iter = it;
  if (info) *info = loc;
  return 0;
}

static int up(drd_ctx* c, double* dst, const double* src, int64_t n);
static int down(drd_ctx* c, double* dst, const double* src, int64_t n);

// ---- learned correction: solver-emitted pairs, training, judged solves -------
static uint64_t splitmix(uint64_t& st) {
  st += 0x9E3779B97F4A7C15ull;
  uint64_t z = st;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// iterations whose iterate a recorded solve keeps: k = round(1.5^j), distinct
static int snap_schedule(int maxit, int* at) {
  int n = 0, last = 0;
  for (double k = 1.0; n < kSnapMax; k *= 1.5) {
    const int ki = (int)std::lround(k);
    if (ki > maxit) break;
    if (ki > last) at[n++] = last = ki;
  }
  return n;
}

// Pairs from the iterates snaps[j] (kept at at[j]) of a solve of A x = b that
// converged in `iters` iterations to x: r = b - A x_k, z = M r (FP32 DCT),
// d = r - A z, e = x - x_k (= A^-1 r); features of (r, d, z) and target
// (e - z) / s_d, the error the DCT solve leaves.
static int learn_pairs(drd_ctx* c, const double* b, const double* x, const int* at, int nsnap, int iters,
                       int* added) {
  Learn* L = c->learn;
  hipStream_t st = c->stream;
  const int64_t n = c->g.nuv();
  *added = 0;
  for (int j = 0; j < nsnap && at[j] <= 0.8 * iters; ++j) {
    const double* xk = L->snaps + (size_t)j * n;
    launch_copy(st, n, b, L->r);
    RCHK(apply_op(c, DRD_OP_VELO2, MV_PLAIN, xk, L->r, -1.0, 1.0, nullptr, nullptr, nullptr));
    velo_prec_apply(st, c->vprec, c->g, L->r, L->z, nullptr);
    launch_copy(st, n, L->r, L->d);
    RCHK(apply_op(c, DRD_OP_VELO2, MV_PLAIN, L->z, L->d, -1.0, 1.0, nullptr, nullptr, nullptr));
    launch_copy(st, n, x, L->e);
    launch_axpby(st, n, -1.0, xk, 1.0, L->e);
    launch_axpby(st, n, -1.0, L->z, 1.0, L->e);
    launch_norms3(st, n, L->r, L->d, L->z, L->ss_part, L->ss);
    if (L->crops == 0) {
      const size_t slot = (size_t)L->next;
      launch_pack_features(st, c->g, L->r, L->d, L->z, c->f[DRD_F_ETA], c->f[DRD_F_BDR_PH_PENALTY], L->ss,
                           L->feat + slot * L->spx * kUnetIn);
      launch_make_target(st, c->g, L->e, L->ss, L->targ + slot * L->spx * kUnetOut);
      L->next = (L->next + 1) % L->cap;
      L->count = std::min(L->count + 1, L->cap);
      ++*added;
      continue;
    }
    launch_pack_features(st, c->g, L->r, L->d, L->z, c->f[DRD_F_ETA], c->f[DRD_F_BDR_PH_PENALTY], L->ss, L->fimg);
    launch_make_target(st, c->g, L->e, L->ss, L->ftarg);
    const int nx = c->g.nx, ny = c->g.ny;
    for (int q = 0; q < L->crops; ++q) {  // random crops of the iterate become pairs
      const size_t slot = (size_t)L->next;
      const int y0 = (int)(splitmix(L->rng) % (uint64_t)(ny - L->ch + 1));
      const int x0 = (int)(splitmix(L->rng) % (uint64_t)(nx - L->cw + 1));
      launch_crop(st, nx, kUnetIn, 2, L->fimg, y0, x0, L->ch, L->cw, L->feat + slot * L->spx * kUnetIn);
      launch_crop(st, nx, kUnetOut, 4, L->ftarg, y0, x0, L->ch, L->cw, L->targ + slot * L->spx * kUnetOut);
      L->next = (L->next + 1) % L->cap;
      L->count = std::min(L->count + 1, L->cap);
      ++*added;
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// BiCGSTAB with the FP32 DCT preconditioner on the BDF2 velocity operator,
// keeping iterates; the solution in x is the solve's own result.
static int learn_recorded_solve(drd_ctx* c, const double* b, double* x, int maxit, double tol,
                                drd_solve_info* info, int* added) {
  int at[kSnapMax];
  const int ns = snap_schedule(maxit, at);
  drd_solve_info loc{};
  RCHK(run_bicgstab(c, DRD_OP_VELO2, DRD_PS_VELO_DCT32, c->g.nuv(), b, x, maxit, tol, &loc, false, at, ns,
                    c->learn->snaps));
  RCHK(learn_pairs(c, b, x, at, ns, loc.iter, added));
  if (info) *info = loc;
  return 0;
}

// The velocity solve of velo_prec = DRD_PREC_LEARNED.  Plain method: BiCGSTAB
// with the FP32 DCT preconditioner (the fastest plain solve).  The first solve
// after drd_learn_enable runs three ways from the same start: FGMRES with the
// correction, FGMRES with the plain preconditioner (the iteration reference)
// and the plain method (the time reference); the correction stays in use only
// if it needs no more iterations than plain FGMRES and no more time than the
// plain method.  A later learned solve that needs more than kLearnSlack times
// the reference iterations is finished by the plain method and the context
// falls back to it.
static double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int velo_solve_learned(drd_ctx* c, const double* b, double* x, int maxit, double tol,
                              drd_solve_info* info) {
  Learn* L = c->learn;
  const int64_t n = c->g.nuv();
  const int op = DRD_OP_VELO2, m = kFgmresRestart;
  hipStream_t st = c->stream;
  auto plain = [&](double* xx, int its, drd_solve_info* out) {
    return run_bicgstab(c, op, DRD_PS_VELO_DCT32, n, b, xx, its, tol, out, false);
  };
  if (L && L->state == DRD_LEARN_JUDGE) {
    launch_copy(st, n, x, L->d);  // x0 for plain FGMRES
    launch_copy(st, n, x, L->z);  // x0 for the plain method
    drd_solve_info li{}, pi{}, bi{};
    HIPCHK(hipStreamSynchronize(st));
    double t0 = now_ms();
    RCHK(run_fgmres(c, op, DRD_PS_VELO_LEARNED, n, b, x, m, maxit, tol, &li));
    HIPCHK(hipStreamSynchronize(st));
    const double tl = now_ms() - t0;
    RCHK(run_fgmres(c, op, DRD_PS_VELO_DCT32, n, b, L->d, m, maxit, tol, &pi));
    HIPCHK(hipStreamSynchronize(st));
    t0 = now_ms();
    RCHK(plain(L->z, maxit, &bi));
    const double tb = now_ms() - t0;
    L->ref_its = pi.iter;
    L->judged_its = li.iter;
    L->last_its = li.iter;
    L->judged_ms[0] = tl;
    L->judged_ms[1] = tb;
    if (li.info == 0 && li.iter <= pi.iter && tl <= tb) {
      L->state = DRD_LEARN_ON;
      *info = li;
    } else {
      L->state = DRD_LEARN_FALLBACK;
      ++L->fallbacks;
      launch_copy(st, n, L->z, x);
      *info = bi;
    }
    return 0;
  }
  if (L && L->state == DRD_LEARN_ON) {
    const int allow = std::min(maxit, std::max(1, (int)std::ceil(kLearnSlack * L->ref_its)));
    launch_copy(st, n, x, L->e);
    drd_solve_info li{};
    RCHK(run_fgmres(c, op, DRD_PS_VELO_LEARNED, n, b, x, m, allow, tol, &li));
    L->last_its = li.iter;
    if (li.info == 0) {
      *info = li;
      return 0;
    }
    // regression: the plain method finishes from the current iterate (from
    // the start if the learned iterate is not finite)
    L->state = DRD_LEARN_FALLBACK;
    ++L->fallbacks;
    if (!std::isfinite(li.resid)) launch_copy(st, n, L->e, x);
    drd_solve_info pi{};
    RCHK(plain(x, std::max(1, maxit - li.iter), &pi));
    pi.iter += li.iter;
    *info = pi;
    return 0;
  }
  return plain(x, maxit, info);
}

static int learn_ready(drd_ctx* c) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  if (!c->learn) return set_err(DRD_E_STATE, "drd_learn_init has not been called");
  return 0;
}

extern "C" int drd_learn_init(drd_ctx* c, int32_t capacity, int32_t crop_h, int32_t crop_w, uint64_t seed) {
  return drd_learn_init_crops(c, capacity, crop_h, crop_w, 0, seed);
}

extern "C" int drd_learn_init_crops(drd_ctx* c, int32_t capacity, int32_t crop_h, int32_t crop_w,
                                    int32_t crops_per_iterate, uint64_t seed) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  if (crops_per_iterate < 0) return set_err(DRD_E_ARG, "crops_per_iterate must be >= 0");
  const Grid& g = c->g;
  if (capacity < 1) return set_err(DRD_E_ARG, "capacity must be positive");
  if (g.nx % 4 || g.ny % 4) return set_err(DRD_E_UNSUPPORTED, "the network needs nx, ny multiples of 4");
  if (crop_h < 4 || crop_w < 4 || crop_h % 4 || crop_w % 4 || crop_h > g.ny || crop_w > g.nx)
    return set_err(DRD_E_ARG, "crop must be multiples of 4 within ny x nx");
  HIPCHK(hipStreamSynchronize(c->stream));
  learn_destroy(c->learn);
  c->learn = nullptr;
  Learn* L = new Learn();
  int err = 0;
  L->net = unet_create(g.ny, g.nx, seed, &err);
  if (L->net) L->tr = unet_create(crop_h, crop_w, seed, &err);
  if (!L->tr) {
    learn_destroy(L);
    return set_err(err ? err : DRD_E_HIP, "network allocation");
  }
  L->ch = crop_h;
  L->cw = crop_w;
  L->cap = capacity;
  L->npx = (int64_t)g.nx * g.ny;
  L->crops = crops_per_iterate;
  L->spx = crops_per_iterate ? (int64_t)crop_h * crop_w : L->npx;
  L->rng = seed ^ 0x6A09E667F3BCC909ull;
  const int64_t nuv = g.nuv(), cpx = (int64_t)crop_h * crop_w;
  auto alloc = [&](void** p, size_t bytes) {
    if (err) return;
    if (hipMalloc(p, bytes) != hipSuccess) {
      err = DRD_E_HIP;
      *p = nullptr;
      return;
    }
    L->allocs.push_back(*p);
  };
  alloc((void**)&L->feat, (size_t)capacity * L->spx * kUnetIn * 2);
  alloc((void**)&L->targ, (size_t)capacity * L->spx * kUnetOut * 4);
  if (L->crops) {
    alloc((void**)&L->fimg, (size_t)L->npx * kUnetIn * 2);
    alloc((void**)&L->ftarg, (size_t)L->npx * kUnetOut * 4);
  }
  alloc((void**)&L->img, (size_t)L->npx * kUnetIn * 2);
  alloc((void**)&L->out, (size_t)L->npx * kUnetOut * 4);
  alloc((void**)&L->cfeat, (size_t)cpx * kUnetIn * 2);
  alloc((void**)&L->ctarg, (size_t)cpx * kUnetOut * 4);
  alloc((void**)&L->ss, 4 * sizeof(double));
  alloc((void**)&L->ss_part, 3 * 256 * sizeof(double));
  alloc((void**)&L->r, nuv * sizeof(double));
  alloc((void**)&L->d, nuv * sizeof(double));
  alloc((void**)&L->z, nuv * sizeof(double));
  alloc((void**)&L->e, nuv * sizeof(double));
  alloc((void**)&L->w, nuv * sizeof(double));
  alloc((void**)&L->dac, 2 * nuv * sizeof(double));
  alloc((void**)&L->part, (2 * (size_t)tiles_of(nuv) + 64) * sizeof(double));
  alloc((void**)&L->dots, 4 * sizeof(double));
  alloc((void**)&L->snaps, (size_t)kSnapMax * nuv * sizeof(double));
  if (err) {
    learn_destroy(L);
    return set_err(DRD_E_HIP, "learned-correction buffers do not fit");
  }
  if (hipMemset(L->img, 0, (size_t)L->npx * kUnetIn * 2) != hipSuccess) {
    learn_destroy(L);
    return set_err(DRD_E_HIP, "hipMemset");
  }
  c->learn = L;
  return 0;
}

extern "C" int drd_learn_collect(drd_ctx* c, const double* b, double tol, double* x, int32_t* pairs_added,
                                 int32_t* iters) {
  RCHK(learn_ready(c));
  if (!b) return set_err(DRD_E_ARG, "null b");
  const int64_t n = c->g.nuv();
  RCHK(up(c, c->rhs, b, n));
  HIPCHK(hipMemsetAsync(c->xv, 0, n * sizeof(double), c->stream));
  int added = 0;
  drd_solve_info info{};
  Learn* L = c->learn;
  const int next0 = L->next, count0 = L->count;
  RCHK(learn_recorded_solve(c, c->rhs, c->xv, std::max(c->p.maxit_velo2, 1), tol, &info, &added));
  RCHK(down(c, x, c->xv, n));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (iters) *iters = info.iter;
  if (info.info != 0) {  // no converged solution, so no error to learn from: the pairs are dropped
    L->next = next0;
    L->count = count0;
    if (pairs_added) *pairs_added = 0;
    return set_err(DRD_E_STATE, "recorded solve did not converge (its pairs are dropped)");
  }
  if (pairs_added) *pairs_added = added;
  return 0;
}

extern "C" int drd_learn_set_every(drd_ctx* c, int32_t every) {
  RCHK(learn_ready(c));
  if (every < 0) return set_err(DRD_E_ARG, "every must be >= 0");
  c->learn->every = every;
  return 0;
}

extern "C" int drd_learn_record(drd_ctx* c, int32_t on) {
  RCHK(learn_ready(c));
  c->learn->record = on != 0;
  return 0;
}

extern "C" int drd_learn_train(drd_ctx* c, int32_t iters, int32_t batch, double lr) {
  RCHK(learn_ready(c));
  Learn* L = c->learn;
  if (L->count == 0) return set_err(DRD_E_STATE, "no pairs collected");
  if (iters < 1 || batch < 1 || !(lr > 0.0)) return set_err(DRD_E_ARG, "iters, batch, lr must be positive");
  hipStream_t st = c->stream;
  HIPCHK(hipStreamSynchronize(st));
  const auto t0 = std::chrono::steady_clock::now();
  const int nx = c->g.nx, ny = c->g.ny;
  double lfirst = 0.0, llast = 0.0;
  for (int it = 0; it < iters; ++it) {
    RCHK(unet_zero_grad(st, L->tr));
    for (int bi = 0; bi < batch; ++bi) {
      const size_t slot = (size_t)(splitmix(L->rng) % (uint64_t)L->count);
      const int wd = L->crops ? L->cw : nx, ht = L->crops ? L->ch : ny;  // a stored crop is taken whole
      const int y0 = (int)(splitmix(L->rng) % (uint64_t)(ht - L->ch + 1));
      const int x0 = (int)(splitmix(L->rng) % (uint64_t)(wd - L->cw + 1));
      launch_crop(st, wd, kUnetIn, 2, L->feat + slot * L->spx * kUnetIn, y0, x0, L->ch, L->cw, L->cfeat);
      launch_crop(st, wd, kUnetOut, 4, L->targ + slot * L->spx * kUnetOut, y0, x0, L->ch, L->cw, L->ctarg);
      RCHK(unet_forward(st, L->tr, L->cfeat, nullptr));
      RCHK(unet_backward(st, L->tr, L->cfeat, L->ctarg));
    }
    const int k = std::max(1, std::min(10, iters / 4));
    if (it < k || it >= iters - k) {
      double loss = 0.0;
      RCHK(unet_read_loss(st, L->tr, &loss));
      if (it < k) lfirst += loss / batch / k;
      if (it >= iters - k) llast += loss / batch / k;
    }
    const double cosf = 0.5 * (1.0 + std::cos(M_PI * it / iters));
    RCHK(unet_adam(st, L->tr, (float)(lr * (0.1 + 0.9 * cosf)), 0.9f, 0.999f, 1e-8f));
  }
  RCHK(unet_copy_params(st, L->net, L->tr));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  L->loss_first = lfirst;
  L->loss_last = llast;
  L->train_iters += iters;
  L->train_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return 0;
}

extern "C" int drd_learn_enable(drd_ctx* c, int32_t on) {
  RCHK(learn_ready(c));
  if (on && c->learn->train_iters == 0) return set_err(DRD_E_STATE, "network not trained");
  c->learn->state = on ? DRD_LEARN_JUDGE : DRD_LEARN_OFF;
  return 0;
}

extern "C" int drd_learn_get_info(drd_ctx* c, drd_learn_info* out) {
  RCHK(learn_ready(c));
  if (!out) return set_err(DRD_E_ARG, "null out");
  const Learn* L = c->learn;
  drd_learn_info o{};
  o.state = L->state;
  o.pairs = L->count;
  o.capacity = L->cap;
  o.train_iters = L->train_iters;
  o.ref_iters = L->ref_its;
  o.judged_iters = L->judged_its;
  o.last_iters = L->last_its;
  o.fallbacks = L->fallbacks;
  o.loss_first = L->loss_first;
  o.loss_last = L->loss_last;
  o.train_ms = L->train_ms;
  o.judged_ms = L->judged_ms[0];
  o.plain_ms = L->judged_ms[1];
  *out = o;
  return 0;
}

extern "C" int drd_learn_get_params(drd_ctx* c, float* params, int32_t n) {
  RCHK(learn_ready(c));
  if (!params) return set_err(DRD_E_ARG, "null params");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (unet_get_params(c->learn->tr, params, n)) return set_err(DRD_E_ARG, "parameter count mismatch");
  return 0;
}

extern "C" int drd_learn_set_params(drd_ctx* c, const float* params, int32_t n) {
  RCHK(learn_ready(c));
  if (!params) return set_err(DRD_E_ARG, "null params");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (unet_set_params(c->learn->tr, params, n) || unet_set_params(c->learn->net, params, n))
    return set_err(DRD_E_ARG, "parameter count mismatch");
  c->learn->train_iters = std::max(c->learn->train_iters, 1);  // loaded parameters count as trained
  return 0;
}

extern "C" int drd_unet_param_count(void) { return unet_param_count(); }

extern "C" int drd_unet_run(int32_t H, int32_t W, const float* params, const uint16_t* in, const float* target,
                            float* out, float* grads, double* loss, float lr, float* params_after) {
  if (!params || !in || !out) return set_err(DRD_E_ARG, "null argument");
  int err = 0;
  UNet* u = unet_create(H, W, 1, &err);
  if (!u) return set_err(err ? err : DRD_E_HIP, "network sizes must be multiples of 4");
  const size_t npx = (size_t)H * W;
  void *din = nullptr, *dout = nullptr, *dtg = nullptr;
  int rc = unet_set_params(u, params, unet_param_count());
  auto ok = [](hipError_t e) { return e == hipSuccess ? 0 : DRD_E_HIP; };
  if (!rc) rc = ok(hipMalloc(&din, npx * kUnetIn * 2));
  if (!rc) rc = ok(hipMalloc(&dout, npx * kUnetOut * 4));
  if (!rc && target) rc = ok(hipMalloc(&dtg, npx * kUnetOut * 4));
  if (!rc) rc = ok(hipMemcpy(din, in, npx * kUnetIn * 2, hipMemcpyHostToDevice));
  if (!rc && target) rc = ok(hipMemcpy(dtg, target, npx * kUnetOut * 4, hipMemcpyHostToDevice));
  if (!rc) rc = unet_forward(0, u, din, (float*)dout);
  if (!rc) rc = ok(hipMemcpy(out, dout, npx * kUnetOut * 4, hipMemcpyDeviceToHost));
  if (!rc && target) {
    // the backward reads the network's own output buffer: run the forward into it
    rc = unet_forward(0, u, din, nullptr);
    if (!rc) rc = unet_zero_grad(0, u);
    if (!rc) rc = unet_backward(0, u, din, (const float*)dtg);
    double l = 0.0;
    if (!rc) rc = unet_read_loss(0, u, &l);
    if (!rc && loss) *loss = l;
    if (!rc && grads) rc = unet_get_grads(u, grads, unet_param_count());
    if (!rc && params_after) {
      rc = unet_adam(0, u, lr, 0.9f, 0.999f, 1e-8f);
      if (!rc) rc = ok(hipDeviceSynchronize());
      if (!rc) rc = unet_get_params(u, params_after, unet_param_count());
    }
  }
  if (!rc) rc = ok(hipDeviceSynchronize());
  hipFree(din);
  hipFree(dout);
  hipFree(dtg);
  unet_destroy(u);
  return rc ? set_err(rc, "network run failed") : 0;
}

// ---- velocity --------------------------------------------------------------
static int velo_prec(const drd_ctx* c, int bdf) {
  switch (c->p.velo_prec) {
    case DRD_PREC_IDENTITY:
      return DRD_PS_IDENTITY;
    case DRD_PREC_JACOBI:
      return DRD_PS_VELO1_JACOBI;
    case DRD_PREC_DCT32:  // the first (BDF1) step takes the same preconditioner built for its operator
      return bdf == 1 ? DRD_PS_VELO1_DCT32 : DRD_PS_VELO_DCT32;
    case DRD_PREC_DCT64:
      return bdf == 1 ? DRD_PS_VELO1_DCT64 : DRD_PS_VELO_DCT64;
    case DRD_PREC_LEARNED:  // plain FP32 DCT unless the correction is judged in (velo_solve_dev)
      return bdf == 1 ? DRD_PS_VELO1_DCT32 : DRD_PS_VELO_DCT32;
    case DRD_PREC_MG:
      return bdf == 1 ? DRD_PS_VELO1_MG : DRD_PS_VELO_MG;
    case DRD_PREC_MG32:
      return bdf == 1 ? DRD_PS_VELO1_MG32 : DRD_PS_VELO_MG32;
    default:  // the reference's: Jacobi for BDF1; identity (TP) or Jacobi (MF) for BDF2
      return bdf == 1 ? DRD_PS_VELO1_JACOBI
                      : (c->p.variant == DRD_VARIANT_MF ? DRD_PS_VELO2_JACOBI : DRD_PS_IDENTITY);
  }
}

// uL(k) TP/VeloSolve_2_combined.f90:272-307, z(0,k) = (k+0.5) dz
static int inflow(drd_ctx* c, double t) {
  const drd_params& p = c->p;
  if (!c->s.lr_d) return 0;  // 'd', or the MF variant (inflow always on)
  const double cf = 1.0 - std::exp(-t / p.twall);
  const double th = p.bdr_upp - p.bdr_low;
  for (int k = 0; k < p.ny; ++k) {
    const double z0 = (k + 0.5) * p.dz;
    if (z0 > p.bdr_low && z0 < p.bdr_upp) {
      const double d = z0 - p.bdr_low - th / 2;
      c->uL_host[k] = p.UU * (-(d * d) / th / th * 4 + 1) * cf;
    } else {
      c->uL_host[k] = 0.0;
    }
  }
  HIPCHK(hipMemcpyAsync(c->uL, c->uL_host, p.ny * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int velo_solve_dev(drd_ctx* c, int bdf, const double* psi1, const double* psi2, const double* u_pre,
                          const double* v_pre, const double* ph_after, const double* rmu_after, double t,
                          const double* pf1, const double* pf2, double* u_after, double* v_after,
                          drd_solve_info* info, bool want_true) {
  if (c->p.xlsB >= 1.0e-8) return set_err(DRD_E_UNSUPPORTED, "slip (xlsB >= 1e-8) branch is empty in the reference");
  RCHK(inflow(c, t));
  VeloRhsArgs a{};
  a.u = c->f[DRD_F_U];
  a.v = c->f[DRD_F_V];
  a.u_pre = u_pre;
  a.v_pre = v_pre;
  a.p = c->f[DRD_F_P];
  a.psi1 = psi1;
  a.psi2 = psi2;
  a.rho = c->f[DRD_F_RHO];
  a.eta = c->f[DRD_F_ETA];
  a.bu = c->f[DRD_F_BDR_PH_USED];
  a.ph_after = ph_after;
  a.rmu_after = rmu_after;
  a.pf1 = pf1 ? pf1 : c->zero_c;
  a.pf2 = pf2 ? pf2 : c->zero_c;
  a.uL = c->uL;
  a.bdr = c->f[DRD_F_BDR_PH];
  if (bdf == 1) {  // unused by the BDF1 expressions
    a.u_pre = c->f[DRD_F_U];
    a.v_pre = c->f[DRD_F_V];
    a.psi1 = a.psi2 = c->zero_c;
  }
  launch_velo_rhs(c->stream, bdf, c->g, c->s, a, c->rhs);
  launch_velo_pack(c->stream, c->g, c->f[DRD_F_U], c->f[DRD_F_V], c->xv);
  if (bdf == 2 && c->guess == 1 && u_pre && v_pre) {  // x0 = 2 x^n - x^{n-1}
    launch_velo_pack(c->stream, c->g, u_pre, v_pre, c->tmp);
    launch_axpby(c->stream, c->g.nuv(), -1.0, c->tmp, 2.0, c->xv);
  }
  const int op = bdf == 1 ? DRD_OP_VELO1 : DRD_OP_VELO2;
  const int maxit = bdf == 1 ? c->p.maxit_velo1 : c->p.maxit_velo2;
  if (bdf == 2 && c->p.velo_prec == DRD_PREC_LEARNED) {
    if (c->learn && c->learn->record) {
      int added = 0;
      RCHK(learn_recorded_solve(c, c->rhs, c->xv, maxit, c->p.tol, info, &added));
    } else {
      RCHK(velo_solve_learned(c, c->rhs, c->xv, maxit, c->p.tol, info));
    }
  } else if (c->p.krylov == DRD_KRYLOV_FGMRES && bdf == 2)
    RCHK(run_fgmres(c, op, velo_prec(c, bdf), c->g.nuv(), c->rhs, c->xv, kFgmresRestart, maxit, c->p.tol, info));
  else
    RCHK(run_bicgstab(c, op, velo_prec(c, bdf), c->g.nuv(), c->rhs, c->xv, maxit, c->p.tol, info, want_true));
  launch_velo_unpack(c->stream, c->g, c->s, c->xv, c->uL, u_after, v_after);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- Cahn-Hilliard ---------------------------------------------------------
static int ph_mu_solve_dev(drd_ctx* c, double* ph_after, double* rmu_after, drd_solve_info* info,
                           bool want_true, bool extrapolate = false) {
  launch_phmu_m0(c->stream, c->g, c->p.phi_mode, c->f[DRD_F_PH], c->f[DRD_F_BDR_PH], c->f[DRD_F_PH_M_0]);
  PhmuRhsArgs a{};
  a.ph = c->f[DRD_F_PH];
  a.rmu = c->f[DRD_F_RMU];
  a.u = c->f[DRD_F_U];
  a.v = c->f[DRD_F_V];
  a.bdr = c->f[DRD_F_BDR_PH];
  a.bu = c->f[DRD_F_BDR_PH_USED];
  a.bg = c->f[DRD_F_BDR_PH_GRAD];
  a.cfB = c->cfB;
  a.cfT = c->cfT;
  a.cosa = c->cosa;
  a.sqrt2 = c->sqrt2;
  launch_phmu_rhs(c->stream, c->g, c->s, a, c->rhs, c->xv, c->pn);
  if (extrapolate) {  // x0 = 2 (phi, mu)^n - (phi, mu)^{n-1}
    launch_phmu_pack(c->stream, c->g, c->f[DRD_F_PH_PRE], c->f[DRD_F_RMU_PRE], c->tmp);
    launch_axpby(c->stream, 2 * c->g.ncell(), -1.0, c->tmp, 2.0, c->xv);
  }
  const int prec = c->p.ch_prec == DRD_CH_PREC_BLOCK    ? DRD_PS_PHMU_BLOCK
                   : c->p.ch_prec == DRD_CH_PREC_RBGS   ? DRD_PS_PHMU_RBGS
                   : c->p.ch_prec == DRD_CH_PREC_RBGS32 ? DRD_PS_PHMU_RBGS32
                                                        : DRD_PS_PHMU_JACOBI;
  RCHK(run_bicgstab(c, DRD_OP_PHMU, prec, 2 * c->g.ncell(), c->rhs, c->xv, c->p.maxit_phmu, c->p.tol, info,
                    want_true));
  launch_phmu_unpack(c->stream, c->g, c->s, c->xv, c->pn, ph_after, rmu_after);
  HIPCHK(hipGetLastError());
  return 0;
}

// ---- pressure ---------------------------------------------------------------
static int presolve_dev(drd_ctx* c, double* psi, double coef, double* p_update, double* mean_div) {
  launch_pressure(c->stream, c->dct, c->g, c->s, c->f[DRD_F_U], c->f[DRD_F_V], c->pre_lcc, coef, psi, p_update,
                  c->dct_work, c->part, c->ks);
  HIPCHK(hipGetLastError());
  if (mean_div) {
    RCHK(read_state(c));
    for (int i = 0; i < c->nb; ++i) c->last_div[i] = c->ks_host[i].red[0] / c->p.nx / c->p.ny;
    *mean_div = c->last_div[0];
  }
  return 0;
}

// ---- host-array entry points ---------------------------------------------
static int up(drd_ctx* c, double* dst, const double* src, int64_t n) {
  if (!src) return 0;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return 0;
}
// a field argument (n its field size): periodic x fills its ghost columns
static int up_field(drd_ctx* c, double* dst, const double* src, int64_t n) {
  RCHK(up(c, dst, src, n));
  if (src) fill_ghosts(c, dst, n);
  return 0;
}
static int down(drd_ctx* c, double* dst, const double* src, int64_t n) {
  if (!dst) return 0;
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  return 0;
}

extern "C" int drd_velosolve_combined(drd_ctx* c, const double* psi, double* u_after, double* v_after,
                                      const double* rmu_after, const double* ph_after, double t,
                                      const double* pf1, const double* pf2, drd_solve_info* info) {
  if (!c || !u_after || !v_after || !rmu_after || !ph_after) return set_err(DRD_E_ARG, "null argument");
  (void)psi;  // temp_p = p (TP/VeloSolve_combined.f90:192)
  const Grid& g = c->g;
  RCHK(up_field(c, c->arg[0], ph_after, g.nc()));
  RCHK(up_field(c, c->arg[1], rmu_after, g.nc()));
  RCHK(up_field(c, c->arg[2], pf1, g.nc()));
  RCHK(up_field(c, c->arg[3], pf2, g.nc()));
  drd_solve_info loc{};
  RCHK(velo_solve_dev(c, 1, nullptr, nullptr, nullptr, nullptr, c->arg[0], c->arg[1], t, pf1 ? c->arg[2] : nullptr,
                      pf2 ? c->arg[3] : nullptr, c->u_after, c->v_after, &loc, true));
  RCHK(down(c, u_after, c->u_after, g.nuf()));
  RCHK(down(c, v_after, c->v_after, g.nvf()));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  return 0;
}

extern "C" int drd_velosolve_2_combined(drd_ctx* c, const double* psi1, const double* psi2, const double* u_pre,
                                        double* u_after, const double* v_pre, double* v_after,
                                        const double* ph_pre, const double* ph_after, const double* rmu_pre,
                                        const double* rmu_after, double t, const double* pf1, const double* pf2,
                                        drd_solve_info* info) {
  if (!c || !psi1 || !psi2 || !u_pre || !u_after || !v_pre || !v_after || !ph_after || !rmu_after)
    return set_err(DRD_E_ARG, "null argument");
  (void)ph_pre;
  (void)rmu_pre;  // not read by the active code
  const Grid& g = c->g;
  RCHK(up_field(c, c->arg[0], ph_after, g.nc()));
  RCHK(up_field(c, c->arg[1], rmu_after, g.nc()));
  RCHK(up_field(c, c->arg[2], pf1, g.nc()));
  RCHK(up_field(c, c->arg[3], pf2, g.nc()));
  RCHK(up_field(c, c->arg[4], psi1, g.nc()));
  RCHK(up_field(c, c->arg[5], psi2, g.nc()));
  RCHK(up_field(c, c->arg[6], u_pre, g.nuf()));
  RCHK(up_field(c, c->arg[7], v_pre, g.nvf()));
  drd_solve_info loc{};
  RCHK(velo_solve_dev(c, 2, c->arg[4], c->arg[5], c->arg[6], c->arg[7], c->arg[0], c->arg[1], t,
                      pf1 ? c->arg[2] : nullptr, pf2 ? c->arg[3] : nullptr, c->u_after, c->v_after, &loc, true));
  RCHK(down(c, u_after, c->u_after, g.nuf()));
  RCHK(down(c, v_after, c->v_after, g.nvf()));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  return 0;
}

extern "C" int drd_presolve_2(drd_ctx* c, double* psi, double coef, double* mean_div) {
  if (!c || !psi) return set_err(DRD_E_ARG, "null argument");
  RCHK(presolve_dev(c, c->arg[0], coef, nullptr, mean_div));
  RCHK(down(c, psi, c->arg[0], c->g.nc()));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int drd_ph_mu_solve(drd_ctx* c, double* ph_after, double* rmu_after, double t, drd_solve_info* info) {
  if (!c || !ph_after || !rmu_after) return set_err(DRD_E_ARG, "null argument");
  (void)t;
  drd_solve_info loc{};
  RCHK(ph_mu_solve_dev(c, c->ph_after, c->rmu_after, &loc, true));
  RCHK(down(c, ph_after, c->ph_after, c->g.nc()));
  RCHK(down(c, rmu_after, c->rmu_after, c->g.nc()));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  return 0;
}

// PhSolve MF/PhSolve.f90:1-126 on the device: RHS, BiCGSTAB(ITER<=500) with
// the DCT-I preconditioner, rmu_after and the phi recovery.
static int ph_solve_dev(drd_ctx* c, double* rmu_after, double* ph_after, drd_solve_info* info, bool want_true) {
  RCHK(aux_ready(c));
  launch_ph_rhs(c->stream, c->g, c->s, c->f[DRD_F_PH], c->f[DRD_F_U], c->f[DRD_F_V], c->f[DRD_F_BDR_PH],
                c->f[DRD_F_BDR_PH_GRAD], c->f[DRD_F_PH_M_INNER], c->f[DRD_F_RMU], c->cosa, c->sqrt2, c->ph_f2,
                c->rhs, c->xv);
  RCHK(run_bicgstab(c, DRD_OP_PH, DRD_PS_PH_DCT, c->g.ncell(), c->rhs, c->xv, c->p.maxit_ph, c->p.tol, info,
                    want_true));
  launch_ph_finish(c->stream, c->g, c->s, c->xv, c->ph_f2, c->f[DRD_F_PH_M_0], rmu_after, ph_after);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int drd_phsolve(drd_ctx* c, double* rmu_after, double* ph_after, double t, drd_solve_info* info) {
  if (!c || !rmu_after || !ph_after) return set_err(DRD_E_ARG, "null argument");
  (void)t;
  drd_solve_info loc{};
  RCHK(ph_solve_dev(c, c->rmu_after, c->ph_after, &loc, true));
  RCHK(down(c, rmu_after, c->rmu_after, c->g.nc()));
  RCHK(down(c, ph_after, c->ph_after, c->g.nc()));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  return 0;
}

extern "C" int drd_matvec(drd_ctx* c, int32_t op, double alpha, const double* x, double beta, double* y) {
  if (!c || !x || !y) return set_err(DRD_E_ARG, "null argument");
  const int64_t n = op_size(c, op);
  if (n < 0) return set_err(DRD_E_ARG, "unknown operator");
  RCHK(up(c, c->tmp, x, n));
  RCHK(up(c, c->tmp2, y, n));
  RCHK(apply_op(c, op, MV_PLAIN, c->tmp, c->tmp2, alpha, beta, nullptr, nullptr, nullptr));
  RCHK(down(c, y, c->tmp2, n));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int drd_psolve(drd_ctx* c, int32_t prec, int32_t op, double* x_out, const double* b_in) {
  if (!c || !x_out || !b_in) return set_err(DRD_E_ARG, "null argument");
  const int64_t n = op_size(c, op);
  if (n < 0) return set_err(DRD_E_ARG, "unknown operator");
  RCHK(up(c, c->tmp, b_in, n));
  RCHK(prec_setup(c, prec));
  const double* res = nullptr;
  RCHK(apply_prec(c, prec, op, c->tmp, c->tmp2, &res, nullptr));
  RCHK(down(c, x_out, res, n));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int drd_bicgstab(drd_ctx* c, int32_t op, int32_t prec, int32_t n, const double* b, double* x,
                            int32_t* iter, double* resid, int32_t* info) {
  if (!c || !b || !x || !iter || !resid || !info) return set_err(DRD_E_ARG, "null argument");
  if (n < 0) {  // LIB/BiCGSTAB.f:157-158
    *info = -1;
    return 0;
  }
  RCHK(up(c, c->rhs, b, n));
  RCHK(up(c, c->xv, x, n));
  drd_solve_info loc{};
  RCHK(run_bicgstab(c, op, prec, n, c->rhs, c->xv, *iter, *resid, &loc, false));
  RCHK(down(c, x, c->xv, n));
  HIPCHK(hipStreamSynchronize(c->stream));
  *iter = loc.iter;
  *resid = loc.resid;
  *info = loc.info;
  return 0;
}

// PhSolve_ori TP/PhSolve_ori.f90:1-234 on the device fields, results into
// ph_out, mu_out (device cell fields)
static int phsolve_ori_dev(drd_ctx* c, double alpha_s, double* ph_out, double* mu_out, drd_solve_info* info) {
  if (c->nb != 1) return set_err(DRD_E_UNSUPPORTED, "PhSolve_ori takes a single-instance context");
  if (c->g.px) return set_err(DRD_E_UNSUPPORTED, "PhSolve_ori runs with Neumann sides");
  RCHK(aux_ready(c));
  const drd_params& p = c->p;
  const double tmp1 = p.I_thickness / p.dz + 0.5 * alpha_s;  // :76-78, :114-116
  c->phori_cb = c->phori_ct = (p.I_thickness / p.dz - 0.5 * alpha_s) / tmp1;
  launch_phori_rhs(c->stream, c->g, c->s, c->f[DRD_F_PH], c->f[DRD_F_U], c->f[DRD_F_V], c->f[DRD_F_ANGLEB],
                   c->f[DRD_F_ANGLET], alpha_s, c->rhs, c->xv, c->pn);
  RCHK(run_bicgstab(c, DRD_OP_PHORI, DRD_PS_PH_DCT, c->g.ncell(), c->rhs, c->xv, 500, p.tol, info, true));  // :148
  launch_phori_finish(c->stream, c->g, c->s, c->phori_cb, c->phori_ct, c->xv, c->pn, ph_out, mu_out);
  HIPCHK(hipGetLastError());
  return 0;
}

extern "C" int drd_phsolve_ori(drd_ctx* c, double alpha_s, double* rmu_after, double* ph_after,
                               drd_solve_info* info) {
  if (!c || !rmu_after || !ph_after) return set_err(DRD_E_ARG, "null argument");
  drd_solve_info loc{};
  RCHK(phsolve_ori_dev(c, alpha_s, c->ph_after, c->rmu_after, &loc));
  RCHK(down(c, ph_after, c->ph_after, c->g.nc()));
  RCHK(down(c, rmu_after, c->rmu_after, c->g.nc()));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  return 0;
}

extern "C" int drd_reinit_solid(drd_ctx* c, int32_t steps, double dt, double alpha_s, drd_solve_info* info) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  if (steps < 0 || !(dt > 0.0)) return set_err(DRD_E_ARG, "steps >= 0 and dt > 0");
  if (c->nb != 1) return set_err(DRD_E_UNSUPPORTED, "the re-initialisation takes a single-instance context");
  if (c->g.px) return set_err(DRD_E_UNSUPPORTED, "the re-initialisation runs with Neumann sides");
  const Grid& g = c->g;
  launch_reinit_start(c->stream, g, c->f[DRD_F_BDR_PH], c->f[DRD_F_PH], c->f[DRD_F_PH_M_0],
                      c->f[DRD_F_PH_M_INNER]);
  const Scal s0 = c->s;
  drd_params q = c->p;
  q.dt = dt;
  c->s = make_scal(q);
  int rc = build_ph_lcc(c, dt);
  drd_solve_info loc{};
  for (int32_t k = 0; k < steps && !rc; ++k) {
    rc = phsolve_ori_dev(c, alpha_s, c->ph_after, c->rmu_after, &loc);
    if (!rc) {
      launch_copy(c->stream, g.nc(), c->rmu_after, c->f[DRD_F_RMU]);
      launch_copy(c->stream, g.nc(), c->ph_after, c->f[DRD_F_PH]);
    }
  }
  c->s = s0;
  const int rc2 = build_ph_lcc(c, c->p.dt);
  RCHK(rc);
  RCHK(rc2);
  launch_reinit_end(c->stream, g, c->f[DRD_F_PH], c->f[DRD_F_BDR_PH], c->f[DRD_F_BDR_PH_PENALTY]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  if (info) *info = loc;
  return 0;
}

// The context drd_bicgstab_cb uses when the caller gives none: Krylov
// workspace for at least n unknowns (a square grid of side m, 2 m^2 >= n),
// made on first use and remade larger when a longer system comes.
static std::mutex g_vec_mu;
static drd_ctx* g_vec_ctx = nullptr;
static int vec_ctx(int64_t n, drd_ctx** out) {
  if (g_vec_ctx && g_vec_ctx->nmax >= n) {
    *out = g_vec_ctx;
    return 0;
  }
  int m = 4;
  while (2 * (int64_t)m * m < n) ++m;
  if (m > 4096) return set_err(DRD_E_ARG, "system too long for one context (n > 2 * 4096^2)");
  if (g_vec_ctx) {
    drd_finalize(g_vec_ctx);
    g_vec_ctx = nullptr;
  }
  drd_params p;
  drd_params_default(&p, m, m);
  RCHK(ctx_create(&p, 1, &g_vec_ctx));
  *out = g_vec_ctx;
  return 0;
}

extern "C" int drd_bicgstab_cb(drd_ctx* ctx, int32_t n, const double* b, double* x, int32_t ldw, int32_t* iter,
                               double* resid, drd_matvec_cb matvec, drd_psolve_cb psolve, int32_t* info) {
  if (!iter || !resid || !info) return set_err(DRD_E_ARG, "null argument");
  *info = 0;
  if (n < 0) {  // LIB/BiCGSTAB.f:155-163
    *info = -1;
    return 0;
  }
  if (ldw < std::max<int32_t>(1, n)) {
    *info = -2;
    return 0;
  }
  if (*iter <= 0) {
    *info = -3;
    return 0;
  }
  if (n == 0) {
    *iter = 0;
    *resid = 0.0;
    return 0;
  }
  if (!b || !x || !matvec) return set_err(DRD_E_ARG, "null argument");
  std::unique_lock<std::mutex> lock(g_vec_mu, std::defer_lock);
  drd_ctx* c = ctx;
  if (!c) {
    lock.lock();
    RCHK(vec_ctx(n, &c));
  }
  if (c->nb != 1) return set_err(DRD_E_ARG, "drd_bicgstab_cb takes a single-instance context");
  if (n > c->nmax) return set_err(DRD_E_ARG, "n exceeds the context's Krylov workspace");
  c->ext_n = n;
  c->ext_mv = matvec;
  c->ext_ps = psolve;
  RCHK(up(c, c->rhs, b, n));
  RCHK(up(c, c->xv, x, n));
  drd_solve_info loc{};
  const int rc = run_bicgstab(c, kOpExt, psolve ? kPsExt : DRD_PS_IDENTITY, n, c->rhs, c->xv, *iter, *resid, &loc,
                              false);
  c->ext_mv = nullptr;
  c->ext_ps = nullptr;
  c->ext_n = 0;
  RCHK(rc);
  RCHK(down(c, x, c->xv, n));
  HIPCHK(hipStreamSynchronize(c->stream));
  *iter = loc.iter;
  *resid = loc.resid;
  *info = loc.info;
  return 0;
}

extern "C" int drd_fgmres(drd_ctx* c, int32_t op, int32_t prec, int32_t n, int32_t restart, const double* b,
                          double* x, int32_t* iter, double* resid, int32_t* info) {
  if (!c || !b || !x || !iter || !resid || !info) return set_err(DRD_E_ARG, "null argument");
  if (n < 0) return set_err(DRD_E_ARG, "negative length");
  RCHK(up(c, c->rhs, b, n));
  RCHK(up(c, c->xv, x, n));
  drd_solve_info loc{};
  RCHK(run_fgmres(c, op, prec, n, c->rhs, c->xv, restart, *iter, *resid, &loc));
  RCHK(down(c, x, c->xv, n));
  HIPCHK(hipStreamSynchronize(c->stream));
  *iter = loc.iter;
  *resid = loc.resid;
  *info = loc.info;
  return 0;
}

// ---- coefficient update ----------------------------------------------------
extern "C" int drd_interpolate_coef(drd_ctx* c, double lambda_eta, const double* px, const double* py,
                                    int32_t np, double radius, const double* sx, const double* sy, int32_t nswm,
                                    double swm_radius) {
  if (!c || np < 0 || nswm < 0 || (np > 0 && (!px || !py)) || (nswm > 0 && (!sx || !sy)))
    return set_err(DRD_E_ARG, "bad argument");
  if (c->nb != 1) return set_err(DRD_E_UNSUPPORTED, "drd_interpolate_coef takes a single-instance context");
  const int64_t npts = (int64_t)np + nswm;
  const int64_t slot = std::max<int64_t>(c->g.nuf(), c->g.nc());  // doubles in an upload slot
  if (2 * npts > slot) return set_err(DRD_E_ARG, "too many particles and swimmers for the grid");
  std::vector<double> pts(2 * (size_t)std::max<int64_t>(npts, 1));
  for (int32_t i = 0; i < np; ++i) {
    pts[2 * (size_t)i] = px[i];
    pts[2 * (size_t)i + 1] = py[i];
  }
  for (int32_t j = 0; j < nswm; ++j) {
    pts[2 * (size_t)(np + j)] = sx[j];
    pts[2 * (size_t)(np + j) + 1] = sy[j];
  }
  double* dev = c->arg[0];  // an upload slot
  if (npts > 0)
    HIPCHK(hipMemcpyAsync(dev, pts.data(), 2 * (size_t)npts * sizeof(double), hipMemcpyHostToDevice, c->stream));
  const drd_params& p = c->p;
  launch_interpolate_coef(c->stream, c->g, p.dx, p.dz, p.I_thickness, p.lambda_rho, lambda_eta, dev, np, radius,
                          nswm, swm_radius, c->f[DRD_F_BDR_PH], c->f[DRD_F_FPD_PH], c->f[DRD_F_RHO],
                          c->f[DRD_F_ETA]);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));  // the host array may be reused on return
  return 0;
}

// ---- particle coupling reductions ----------------------------------------------
// Bodies (x, y, radius, fx, fy) go to upload slot 0; the per-chunk sums to slot
// 1 and the per-body results to slot 2.
static int upload_bodies(drd_ctx* c, const double* bx, const double* by, const double* br, const double* fx,
                         const double* fy, int32_t nbody, double lx, double ly, int* chunks) {
  if (!c || nbody < 0 || (nbody > 0 && (!bx || !by || !br)) || !(lx > 0.0) || !(ly > 0.0))
    return set_err(DRD_E_ARG, "bad argument");
  if (c->nb != 1) return set_err(DRD_E_UNSUPPORTED, "particle coupling takes a single-instance context");
  const int64_t slot = std::max<int64_t>(c->g.nuf(), c->g.nc());
  if (5 * (int64_t)nbody > slot || nbody > 65535) return set_err(DRD_E_ARG, "too many bodies for the grid");
  std::vector<double> h(5 * (size_t)std::max(nbody, 1));
  for (int32_t b = 0; b < nbody; ++b) {
    double* e = &h[5 * (size_t)b];
    e[0] = bx[b];
    e[1] = by[b];
    e[2] = br[b];
    e[3] = fx ? fx[b] : 0.0;
    e[4] = fy ? fy[b] : 0.0;
  }
  if (nbody > 0)
    HIPCHK(hipMemcpyAsync(c->arg[0], h.data(), 5 * (size_t)nbody * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  *chunks = body_chunks(c->g, nbody, slot);
  HIPCHK(hipStreamSynchronize(c->stream));  // h is released on return
  return 0;
}

extern "C" int drd_particle_force(drd_ctx* c, const double* bx, const double* by, const double* br,
                                  const double* fx, const double* fy, int32_t nbody, double lx, double ly) {
  if (nbody > 0 && (!fx || !fy)) return set_err(DRD_E_ARG, "bad argument");
  int chunks = 1;
  RCHK(upload_bodies(c, bx, by, br, fx, fy, nbody, lx, ly, &chunks));
  const drd_params& p = c->p;
  launch_body_sums(c->stream, c->g, p.dx, p.dz, p.I_thickness, lx, ly, c->arg[0], nbody, chunks, nullptr, nullptr,
                   c->arg[1], c->arg[2]);
  launch_body_spread(c->stream, c->g, p.dx, p.dz, p.I_thickness, lx, ly, c->arg[0], nbody, c->arg[2],
                     c->f[DRD_F_PFORCE1], c->f[DRD_F_PFORCE2]);
  fill_ghosts(c, c->f[DRD_F_PFORCE1], c->g.nc());  // periodic x: the ghost columns the momentum rows read
  fill_ghosts(c, c->f[DRD_F_PFORCE2], c->g.nc());
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

extern "C" int drd_particle_velocity(drd_ctx* c, const double* bx, const double* by, const double* br,
                                     int32_t nbody, double lx, double ly, double* u_aver, double* v_aver) {
  if (nbody > 0 && (!u_aver || !v_aver)) return set_err(DRD_E_ARG, "bad argument");
  int chunks = 1;
  RCHK(upload_bodies(c, bx, by, br, nullptr, nullptr, nbody, lx, ly, &chunks));
  if (nbody == 0) return 0;
  const drd_params& p = c->p;
  launch_body_sums(c->stream, c->g, p.dx, p.dz, p.I_thickness, lx, ly, c->arg[0], nbody, chunks, c->f[DRD_F_U],
                   c->f[DRD_F_V], c->arg[1], c->arg[2]);
  HIPCHK(hipGetLastError());
  std::vector<double> fin(3 * (size_t)nbody);
  HIPCHK(hipMemcpyAsync(fin.data(), c->arg[2], fin.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int32_t b = 0; b < nbody; ++b) {
    u_aver[b] = fin[3 * (size_t)b + 1];
    v_aver[b] = fin[3 * (size_t)b + 2];
  }
  return 0;
}

// ---- measurement hook ---------------------------------------------------------
// launches per timed group of the instrumented kernel
static int tm_group(const drd_ctx* c) { return c->tm_which == DRD_K_PHMU_SWEEP32 ? kRbgsMiddle : 1; }

extern "C" int drd_kernel_timer(drd_ctx* c, int32_t which, int32_t capacity) {
  if (!c || capacity < 0) return set_err(DRD_E_ARG, "bad argument");
  if (capacity > 0 && which != DRD_K_VELO_MATVEC && which != DRD_K_PHMU_SWEEP32)
    return set_err(DRD_E_UNSUPPORTED, "kernel not instrumented");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto e : c->tm_ev) HIPCHK(hipEventDestroy(e));
  c->tm_ev.clear();
  c->tm_n = 0;
  c->tm_which = capacity > 0 ? which : -1;
  c->tm_ev.resize(2 * (size_t)capacity);
  for (auto& e : c->tm_ev) HIPCHK(hipEventCreate(&e));
  return 0;
}

extern "C" int drd_kernel_timer_read(drd_ctx* c, int32_t* count, double* total_ms, double* bytes) {
  if (!c || !count || !total_ms || !bytes) return set_err(DRD_E_ARG, "bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  double tot = 0.0;
  for (int i = 0; i < c->tm_n; ++i) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->tm_ev[2 * i], c->tm_ev[2 * i + 1]));
    tot += ms;
  }
  *count = c->tm_n * tm_group(c);
  *total_ms = tot;
  if (c->tm_which == DRD_K_PHMU_SWEEP32)  // x in and out, b: 2 floats per cell each; bu, M_inner, M_0: 1 each
    *bytes = 4.0 * 9.0 * (double)c->g.ncell();
  else  // x, y, rtld: Nuv doubles each; rho, eta, chi: C doubles each
    *bytes = 8.0 * (3.0 * (double)c->g.nuv() + 3.0 * (double)c->g.ncell());
  return 0;
}

extern "C" int drd_kernel_timer_launches(drd_ctx* c, int32_t capacity, float* ms, int32_t* count) {
  if (!c || !count || capacity < 0 || (capacity > 0 && !ms)) return set_err(DRD_E_ARG, "bad argument");
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->tm_n && i < capacity; ++i) {
    HIPCHK(hipEventElapsedTime(&ms[i], c->tm_ev[2 * i], c->tm_ev[2 * i + 1]));
    ms[i] /= (float)tm_group(c);
  }
  *count = c->tm_n;
  return 0;
}

extern "C" int drd_bench_kernel(drd_ctx* c, int32_t which, int32_t reps, double* avg_ms, double* bytes) {
  if (!c || !avg_ms || !bytes || reps <= 0) return set_err(DRD_E_ARG, "bad argument");
  const Grid& g = c->g;
  const double C8 = 8.0 * (double)g.ncell(), N8 = 8.0 * (double)g.nuv();
  KState h{};
  h.tol = 0.0;
  h.maxit = 1 << 30;
  h.iter = 2;
  h.rho = h.rho1 = h.alpha = h.omega = 1.0;
  h.bnrm2 = 1.0;
  *c->ks_host = h;
  HIPCHK(hipMemcpyAsync(c->ks, c->ks_host, sizeof(KState), hipMemcpyHostToDevice, c->stream));
  const Coef k = coef_of(c);
  auto launch = [&]() {
    switch (which) {
      case DRD_K_VELO_MATVEC:
        launch_velo_matvec(c->stream, 2, MV_DOT1, g, c->s, k, c->kw[2], c->kw[3], 1.0, 0.0, c->kw[1], nullptr,
                           c->part, c->ks, false);
        break;
      case DRD_K_PHMU_MATVEC:
        launch_phmu_matvec(c->stream, MV_DOT1, g, c->s, k, c->kw[2], c->kw[3], 1.0, 0.0, c->kw[1], nullptr,
                           c->part, c->ks, false);
        break;
      case DRD_K_BICG_XR:
        launch_bicg_xr(c->stream, g.nuv(), 1, c->xv, c->kw[0], c->kw[1], c->kw[5], c->kw[6], c->kw[4], c->part,
                       c->ks, false);
        break;
      case DRD_K_PRESSURE:
        launch_pressure(c->stream, c->dct, g, c->s, c->f[DRD_F_U], c->f[DRD_F_V], c->pre_lcc, 1.0, c->arg[0],
                        nullptr, c->dct_work, c->part, nullptr);
        break;
      case DRD_K_COPY:
        launch_copy(c->stream, g.nuv(), c->kw[2], c->kw[3]);
        break;
      case DRD_K_UNET_FWD:
        (void)unet_forward(c->stream, c->learn->net, c->learn->img, c->learn->out);
        break;
      case DRD_K_VELO_PREC:
        velo_prec_apply(c->stream, c->vprec, g, c->kw[2], c->kw[3], nullptr);
        break;
      case DRD_K_VELO_MG:
        velo_mg_apply(c->stream, c->vmg, c->kw[2], c->kw[3], nullptr);
        break;
      case DRD_K_PHMU_SWEEP32: {  // the first iterate -> the second, with the FP32 b
        const float* wa = (const float*)c->rb_work;
        launch_phmu_rb_sweep32(c->stream, g, c->s, phmu_c32(c), wa + 4 * g.ncell(), wa,
                               (float*)c->rb_work + 2 * g.ncell(), c->ks);
        break;
      }
    }
  };
  switch (which) {
    case DRD_K_VELO_MATVEC:  // x, y, rtld (Nuv each) + rho, eta, chi
      *bytes = 3 * N8 + 3 * C8;
      break;
    case DRD_K_PHMU_MATVEC:  // x, y, rtld (2C each) + bdr_used, M_inner, M_0
      *bytes = 6 * C8 + 3 * C8;
      break;
    case DRD_K_BICG_XR:  // x rw, r rw, rtld, phat, shat, t
      *bytes = 8 * N8;
      break;
    case DRD_K_PRESSURE:  // u, v in; psi out (p untouched here)
      *bytes = N8 + 2 * C8 + C8;
      break;
    case DRD_K_COPY:
      *bytes = 2 * N8;
      break;
    case DRD_K_UNET_FWD:  // MFMA FLOPs, not bytes
      if (!c->learn) return set_err(DRD_E_STATE, "drd_learn_init has not been called");
      *bytes = unet_forward_flops(c->learn->net);
      break;
    case DRD_K_VELO_PREC:  // r in, z out (FP64), scaling in (FP32 or FP64)
      if (!c->vprec) return set_err(DRD_E_STATE, "no DCT velocity preconditioner in this context");
      RCHK(prec_setup(c, c->vprec_fp32 ? DRD_PS_VELO_DCT32 : DRD_PS_VELO_DCT64));
      *bytes = 2 * N8 + (c->vprec_fp32 ? N8 / 2 : N8);
      break;
    case DRD_K_VELO_MG:  // r in, z out (FP64); the context's hierarchy, else the one its velo_prec names
      RCHK(prec_setup(c, (c->vmg ? velo_mg_f32(c->vmg) : c->p.velo_prec == DRD_PREC_MG32) ? DRD_PS_VELO_MG32
                                                                                        : DRD_PS_VELO_MG));
      *bytes = 2 * N8;
      break;
    case DRD_K_PHMU_SWEEP32: {  // x in and out, b: 2 floats per cell each; bu, M_inner, M_0: 1 each
      RCHK(prec_setup(c, DRD_PS_PHMU_RBGS32));
      const double* res = nullptr;  // one application fills the iterates and the FP32 b
      RCHK(apply_prec(c, DRD_PS_PHMU_RBGS32, DRD_OP_PHMU, c->kw[2], c->kw[3], &res, nullptr));
      *bytes = 4.0 * 9.0 * (double)g.ncell();
      break;
    }
    default:
      return set_err(DRD_E_ARG, "unknown kernel id");
  }
  launch();  // warm
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) launch();
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *avg_ms = (double)ms / reps;
  return 0;
}

// ---- time step (TP/mcl6.f90:257-303 and :395-502; MF/mcl6.f90:239-300, :388-440) --
// One step of every instance of the context's group; per-instance reports
// go to out[0 .. nb).
static int step_dev(drd_ctx* c, int bdf, double t, drd_step_info* out) {
  if (bdf != 1 && bdf != 2) return set_err(DRD_E_ARG, "bdf must be 1 or 2");
  const drd_params& p = c->p;
  drd_step_info loc{};
  const double gamma1 = std::min(1.0, p.lambda_rho);
  double** F = c->f;
  if (p.variant == DRD_VARIANT_MF)  // MF/mcl6.f90:239, :388
    RCHK(ph_solve_dev(c, c->rmu_after, c->ph_after, &loc.phmu, false));
  else
    RCHK(ph_mu_solve_dev(c, c->ph_after, c->rmu_after, &loc.phmu, false, bdf == 2 && c->guess == 1));
  for (int i = 0; i < c->nb; ++i) out[i].phmu = c->nb == 1 ? loc.phmu : c->last[i];
  if (bdf == 1) {
    RCHK(velo_solve_dev(c, 1, nullptr, nullptr, nullptr, nullptr, c->ph_after, c->rmu_after, t + p.dt,
                        F[DRD_F_PFORCE1], F[DRD_F_PFORCE2], c->u_after, c->v_after, &loc.velo, false));
  } else {
    RCHK(velo_solve_dev(c, 2, F[DRD_F_PSI1], F[DRD_F_PSI2], F[DRD_F_U_PRE], F[DRD_F_V_PRE], c->ph_after,
                        c->rmu_after, t + p.dt, F[DRD_F_PFORCE1], F[DRD_F_PFORCE2], c->u_after, c->v_after,
                        &loc.velo, false));
  }
  for (int i = 0; i < c->nb; ++i) out[i].velo = c->nb == 1 ? loc.velo : c->last[i];
  // u_pre = u; u = u_after (pointer rotation: u_after is fully rewritten next step)
  std::swap(F[DRD_F_U_PRE], F[DRD_F_U]);
  std::swap(F[DRD_F_U], c->u_after);
  std::swap(F[DRD_F_V_PRE], F[DRD_F_V]);
  std::swap(F[DRD_F_V], c->v_after);
  // ph_pre = ph; ph = ph_after; rmu likewise
  std::swap(F[DRD_F_PH_PRE], F[DRD_F_PH]);
  std::swap(F[DRD_F_PH], c->ph_after);
  std::swap(F[DRD_F_RMU_PRE], F[DRD_F_RMU]);
  std::swap(F[DRD_F_RMU], c->rmu_after);
  double coef;
  if (bdf == 1) {
    coef = p.re * gamma1 / p.dt;
  } else {
    std::swap(F[DRD_F_PSI1], F[DRD_F_PSI2]);  // psi1 = psi2 (psi2 fully rewritten below)
    coef = 3 * p.re * gamma1 / 2. / p.dt;
  }
  RCHK(presolve_dev(c, F[DRD_F_PSI2], coef, F[DRD_F_P], &loc.div_before));
  for (int i = 0; i < c->nb; ++i) out[i].div_before = c->last_div[i];
  return 0;
}

extern "C" int drd_step(drd_ctx* c, int32_t bdf, double t, drd_step_info* info) {
  if (!c) return set_err(DRD_E_ARG, "null ctx");
  drd_step_info loc{};
  RCHK(step_dev(c, bdf, t, &loc));
  if (info) *info = loc;
  return 0;
}

// ---- snapshots in the reference's format (TP/inout.f90:1-61) -----------------
// Fortran unformatted sequential records (4-byte length markers around each
// record), every array written as ((a(i,k), k=0,ny), i=0,nx):
//   x, y (cell centres (i+0.5) dx, (k+0.5) dz), u, v, p, f, f2, t.
static void put_record(std::ofstream& os, const double* v, size_t n) {
  const uint32_t len = (uint32_t)(n * sizeof(double));
  os.write(reinterpret_cast<const char*>(&len), 4);
  os.write(reinterpret_cast<const char*>(v), (std::streamsize)len);
  os.write(reinterpret_cast<const char*>(&len), 4);
}
static bool get_record(std::ifstream& is, double* v, size_t n) {
  uint32_t a = 0, b = 0;
  is.read(reinterpret_cast<char*>(&a), 4);
  if (!is || a != n * sizeof(double)) return false;
  is.read(reinterpret_cast<char*>(v), (std::streamsize)a);
  is.read(reinterpret_cast<char*>(&b), 4);
  return (bool)is && a == b;
}

// element (i, k) of a field stored with its ghost layout
static int64_t field_index(const drd_params& p, int32_t fid, int i, int k) {
  if (fid == DRD_F_U || fid == DRD_F_U_PRE) return (int64_t)(i + 1) + (int64_t)(p.nx + 3) * (k + 1);
  if (fid == DRD_F_V || fid == DRD_F_V_PRE) return (int64_t)(i + 1) + (int64_t)(p.nx + 2) * k;
  return (int64_t)(i + 1) + (int64_t)(p.nx + 2) * (k + 1);
}

static bool cell_field(int32_t fid) {
  return fid >= 0 && fid < DRD_F_COUNT && fid != DRD_F_U && fid != DRD_F_V && fid != DRD_F_U_PRE &&
         fid != DRD_F_V_PRE && fid != DRD_F_ANGLEB && fid != DRD_F_ANGLET;
}

extern "C" int drd_write_snapshot(drd_ctx* c, const char* path, int32_t f_field, int32_t f2_field, double t) {
  if (!c || !path) return set_err(DRD_E_ARG, "null argument");
  if (!cell_field(f_field) || !cell_field(f2_field)) return set_err(DRD_E_ARG, "f, f2 must be cell fields");
  const drd_params& p = c->p;
  const int M = p.nx, N = p.ny;
  const size_t n = (size_t)(M + 1) * (N + 1);
  std::vector<double> rec(n), host;
  std::ofstream os(path, std::ios::binary);
  if (!os) return set_err(DRD_E_ARG, std::string("cannot open ") + path);
  auto emit = [&](auto value) {
    size_t q = 0;
    for (int i = 0; i <= M; ++i)
      for (int k = 0; k <= N; ++k) rec[q++] = value(i, k);
    put_record(os, rec.data(), n);
  };
  emit([&](int i, int) { return (i + 0.5) * p.dx; });
  emit([&](int, int k) { return (k + 0.5) * p.dz; });
  for (int32_t fid : {(int32_t)DRD_F_U, (int32_t)DRD_F_V, (int32_t)DRD_F_P, f_field, f2_field}) {
    host.resize((size_t)drd_field_size(&p, fid));
    RCHK(drd_get_field(c, fid, host.data()));
    emit([&](int i, int k) { return host[(size_t)field_index(p, fid, i, k)]; });
  }
  put_record(os, &t, 1);
  os.flush();
  if (!os) return set_err(DRD_E_HIP, std::string("write failed: ") + path);
  return 0;
}

// Reads all seven arrays and t (the reference's input reads six: it skips
// f2, TP/inout.f90:41-61) and sets u, v, p, f, f2 at their (0:nx, 0:ny)
// elements; the other ghost entries keep their values.
extern "C" int drd_read_snapshot(drd_ctx* c, const char* path, int32_t f_field, int32_t f2_field, double* t) {
  if (!c || !path) return set_err(DRD_E_ARG, "null argument");
  if (!cell_field(f_field) || !cell_field(f2_field)) return set_err(DRD_E_ARG, "f, f2 must be cell fields");
  const drd_params& p = c->p;
  const int M = p.nx, N = p.ny;
  const size_t n = (size_t)(M + 1) * (N + 1);
  std::vector<double> rec(n), host;
  std::ifstream is(path, std::ios::binary);
  if (!is) return set_err(DRD_E_ARG, std::string("cannot open ") + path);
  for (int r = 0; r < 2; ++r)
    if (!get_record(is, rec.data(), n)) return set_err(DRD_E_ARG, "not a snapshot of this grid size");
  for (int32_t fid : {(int32_t)DRD_F_U, (int32_t)DRD_F_V, (int32_t)DRD_F_P, f_field, f2_field}) {
    if (!get_record(is, rec.data(), n)) return set_err(DRD_E_ARG, "truncated snapshot");
    host.resize((size_t)drd_field_size(&p, fid));
    RCHK(drd_get_field(c, fid, host.data()));
    size_t q = 0;
    for (int i = 0; i <= M; ++i)
      for (int k = 0; k <= N; ++k) host[(size_t)field_index(p, fid, i, k)] = rec[q++];
    RCHK(drd_set_field(c, fid, host.data()));
  }
  double tv = 0.0;
  if (!get_record(is, &tv, 1)) return set_err(DRD_E_ARG, "truncated snapshot");
  if (t) *t = tv;
  return 0;
}

// ---- batches: groups of equal-size instances stepped by batched launches -----
struct drd_batch {
  std::vector<drd_ctx*> groups;
  std::vector<std::pair<int, int>> where;  // instance -> (group, index in the group)
  std::vector<std::vector<drd_step_info>> infos;
};

extern "C" int drd_batch_destroy(drd_batch* b) {
  if (!b) return 0;
  for (drd_ctx* c : b->groups) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    ctx_free(c);
    delete c;
  }
  delete b;
  return 0;
}

extern "C" int drd_batch_create(int32_t count, const drd_params* params, drd_batch** out) {
  if (count < 1 || !params || !out) return set_err(DRD_E_ARG, "bad argument");
  std::vector<int> gid(count, -1);
  std::vector<int> rep_of;  // group -> first instance
  for (int i = 0; i < count; ++i) {
    const drd_params& q = params[i];
    if (q.velo_prec != DRD_PREC_DCT32 && q.velo_prec != DRD_PREC_DCT64 && q.velo_prec != DRD_PREC_JACOBI &&
        q.velo_prec != DRD_PREC_IDENTITY && q.velo_prec != DRD_PREC_REFERENCE && q.velo_prec != DRD_PREC_MG &&
        q.velo_prec != DRD_PREC_MG32)
      return set_err(DRD_E_UNSUPPORTED,
                     "batched steps take the DCT, multigrid, Jacobi or identity velocity preconditioner");
    if (q.krylov != DRD_KRYLOV_BICGSTAB) return set_err(DRD_E_UNSUPPORTED, "batched steps use BiCGSTAB");
    if (q.lr_bd_type == 'p') return set_err(DRD_E_UNSUPPORTED, "periodic x takes a single-instance context");
    for (size_t gI = 0; gI < rep_of.size(); ++gI) {
      const drd_params& r = params[rep_of[gI]];
      if (r.nx == q.nx && r.ny == q.ny) {
        if (std::memcmp(&r, &q, sizeof(drd_params)) != 0)
          return set_err(DRD_E_ARG,
                         "instances of one grid size must share every parameter (put per-instance penalties "
                         "into bdr_ph_penalty)");
        gid[i] = (int)gI;
        break;
      }
    }
    if (gid[i] < 0) {
      gid[i] = (int)rep_of.size();
      rep_of.push_back(i);
    }
  }
  drd_batch* b = new drd_batch();
  std::vector<int> counts(rep_of.size(), 0);
  b->where.resize(count);
  for (int i = 0; i < count; ++i) b->where[i] = {gid[i], counts[gid[i]]++};
  for (size_t gI = 0; gI < rep_of.size(); ++gI) {
    drd_ctx* c = nullptr;
    const int rc = ctx_create(&params[rep_of[gI]], counts[gI], &c);
    if (rc) {
      drd_batch_destroy(b);
      return rc;
    }
    b->groups.push_back(c);  // captured chunks are opt-in (drd_batch_use_graphs)
    b->infos.emplace_back(counts[gI]);
  }
  *out = b;
  return 0;
}

static int batch_locate(drd_batch* b, int32_t inst, int32_t field, drd_ctx** c, char** dev, int64_t* n) {
  if (!b || inst < 0 || inst >= (int32_t)b->where.size()) return set_err(DRD_E_ARG, "bad instance");
  *c = b->groups[b->where[inst].first];
  if (field < 0 || field >= DRD_F_COUNT) return set_err(DRD_E_ARG, "bad field id");
  *n = drd_field_size(&(*c)->p, field);
  *dev = reinterpret_cast<char*>((*c)->f[field]) + (int64_t)b->where[inst].second * (*c)->bs;
  return 0;
}

extern "C" int drd_batch_set_field(drd_batch* b, int32_t inst, int32_t field, const double* host) {
  drd_ctx* c = nullptr;
  char* dev = nullptr;
  int64_t n = 0;
  RCHK(batch_locate(b, inst, field, &c, &dev, &n));
  if (!host) return set_err(DRD_E_ARG, "null host array");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(dev, host, n * sizeof(double), hipMemcpyHostToDevice));
  if (field == DRD_F_ANGLEB || field == DRD_F_ANGLET) RCHK(upload_angles(c, b->where[inst].second));
  return 0;
}

extern "C" int drd_batch_get_field(drd_batch* b, int32_t inst, int32_t field, double* host) {
  drd_ctx* c = nullptr;
  char* dev = nullptr;
  int64_t n = 0;
  RCHK(batch_locate(b, inst, field, &c, &dev, &n));
  if (!host) return set_err(DRD_E_ARG, "null host array");
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(host, dev, n * sizeof(double), hipMemcpyDeviceToHost));
  return 0;
}

// The tile autotuner: times the velocity preconditioner's transpose tile
// shapes on the context's current fields (all instances of its group) and
// keeps the fastest; captured graphs are dropped (their kernels change).
static int tune_ctx(drd_ctx* c, int32_t* shape, double* ms) {
  if (!c->vprec) {
    *shape = -1;
    return 0;
  }
  RCHK(prec_setup(c, c->vprec_fp32 ? DRD_PS_VELO_DCT32 : DRD_PS_VELO_DCT64));
  const int r = velo_prec_tune(c->stream, c->vprec, c->g, c->kw[0], c->kw[1], 5, ms);
  if (r < 0) return set_err(DRD_E_HIP, "autotune failed");
  HIPCHK(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  *shape = r;
  return 0;
}

extern "C" int drd_tune(drd_ctx* c, int32_t* shape, double* ms) {
  if (!c || !shape || !ms) return set_err(DRD_E_ARG, "null argument");
  return tune_ctx(c, shape, ms);
}

extern "C" int drd_batch_tune(drd_batch* b, int32_t* shapes, double* ms, int32_t cap) {
  if (!b || !shapes || !ms) return set_err(DRD_E_ARG, "null argument");
  for (int gI = 0; gI < (int)b->groups.size() && gI < cap; ++gI) {
    BatchScope scope(b->groups[gI]);
    RCHK(tune_ctx(b->groups[gI], &shapes[gI], &ms[3 * gI]));
  }
  return 0;
}

extern "C" int drd_batch_set_initial_guess(drd_batch* b, int32_t mode) {
  if (!b) return set_err(DRD_E_ARG, "null batch");
  for (drd_ctx* c : b->groups) RCHK(drd_set_initial_guess(c, mode));
  return 0;
}

extern "C" int drd_batch_use_graphs(drd_batch* b, int32_t on) {
  if (!b) return set_err(DRD_E_ARG, "null batch");
  for (drd_ctx* c : b->groups) RCHK(drd_use_graphs(c, on));
  return 0;
}

extern "C" int drd_batch_groups(drd_batch* b, int32_t* ngroups, int32_t* desc, int32_t cap) {
  if (!b || !ngroups) return set_err(DRD_E_ARG, "null argument");
  *ngroups = (int32_t)b->groups.size();
  for (int gI = 0; gI < (int)b->groups.size() && desc && gI < cap; ++gI) {
    desc[3 * gI] = b->groups[gI]->p.nx;
    desc[3 * gI + 1] = b->groups[gI]->p.ny;
    desc[3 * gI + 2] = b->groups[gI]->nb;
  }
  return 0;
}

// The size groups step at once, each from a host thread of its own on its
// own stream (a group's host polls between iteration chunks then overlap the
// other groups' kernels); the first group's error, if any, is reported.
// Groups that capture iteration chunks as graphs step one after the other:
// concurrent stream captures from several threads fail in the runtime.
extern "C" int drd_batch_step(drd_batch* b, int32_t bdf, double t, drd_step_info* infos) {
  if (!b) return set_err(DRD_E_ARG, "null batch");
  const size_t G = b->groups.size();
  std::vector<int> rc(G, 0);
  std::vector<std::string> msg(G);
  auto run = [&](size_t gI) {
    drd_ctx* c = b->groups[gI];
    BatchScope scope(c);
    rc[gI] = step_dev(c, bdf, t, b->infos[gI].data());
    if (rc[gI]) msg[gI] = g_err;
  };
  bool graphs = false;
  for (drd_ctx* c : b->groups) graphs = graphs || c->use_graphs;
  if (G == 1 || graphs) {
    for (size_t gI = 0; gI < G; ++gI) run(gI);
  } else {
    std::vector<std::thread> th;
    size_t started = 0;
    try {
      th.reserve(G);
      for (; started < G; ++started) th.emplace_back(run, started);
    } catch (...) {  // no thread for the rest (system limits): they step on this one
    }
    for (size_t gI = started; gI < G; ++gI) run(gI);
    for (auto& x : th) x.join();
  }
  for (size_t gI = 0; gI < G; ++gI)
    if (rc[gI]) return set_err(rc[gI], msg[gI]);
  if (infos)
    for (size_t i = 0; i < b->where.size(); ++i) infos[i] = b->infos[b->where[i].first][b->where[i].second];
  return 0;
}
