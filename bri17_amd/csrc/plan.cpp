// Host-side planning of the periodic stiffness operator.
//
// The reference builds one FFTW c2c plan per component at construction
// (tests/test_bri17.cpp:110-128) and recomputes the modal stiffness with libm
// trigonometry at every frequency (bri17.hpp:255-264).  Here a plan holds, per
// grid, device-resident twiddle tables and per-axis modal coefficient tables
// (phi, chi, psi) evaluated once in extended precision on the host, plus the
// launch geometry of the three kernels.  Tables depend on (shape, L) only;
// mu and nu enter the kernels as scalars.
#include "plan.hpp"

#include "kernels_static.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <sstream>

namespace b17 {

// LDS per plane workgroup aimed for when choosing the middle-axis split
// (rows plus twiddle tables): four such workgroups fit the 160 KiB of a CU.
constexpr size_t kPlaneLds = 40 * 1024;

static thread_local std::string t_err;
void set_error(const std::string& msg) { t_err = msg; }
const std::string& get_error() { return t_err; }

static Status fail(int code, const std::string& m) { return Status{code, m}; }

std::vector<int> default_radices(int n) {
  std::vector<int> r;
  if (n < 1) return r;
  while (n % 16 == 0) r.push_back(16), n /= 16;
  if (n % 8 == 0) r.push_back(8), n /= 8;
  if (n % 4 == 0) r.push_back(4), n /= 4;
  if (n % 2 == 0) r.push_back(2), n /= 2;
  while (n % 9 == 0) r.push_back(9), n /= 9;
  for (int p : {3, 5, 7, 11, 13})
    while (n % p == 0) r.push_back(p), n /= p;
  // larger prime factors: one direct-DFT pass each
  for (int p = 17; (long long)p * p <= n; p += 2)
    while (n % p == 0) r.push_back(p), n /= p;
  if (n > 1) r.push_back(n);
  return r;
}

static bool is_prime(int n) {
  if (n < 2) return false;
  for (int p = 2; (long long)p * p <= n; p++)
    if (n % p == 0) return false;
  return true;
}

bool make_fft_desc(int n, int tws, const std::vector<int>& radices, FftDesc& d,
                   std::string* why, bool want_wave) {
  std::memset(&d, 0, sizeof(d));
  d.n = n;
  d.tws = tws;
  if (n == 1) {
    d.npass = 0;
    d.tpt = 1;
    d.wave = 0;
    return true;
  }
  int prod = 1, rmax = 1;
  for (int r : radices) {
    if (!(r == 2 || r == 3 || r == 4 || r == 5 || r == 7 || r == 8 || r == 9 || r == 11 ||
          r == 13 || r == 16 || (r > 16 && is_prime(r)))) {
      if (why) *why = "unsupported radix " + std::to_string(r);
      return false;
    }
    prod *= r;
    rmax = std::max(rmax, r);
  }
  if (prod != n || radices.empty() || (int)radices.size() > kMaxPass) {
    if (why) *why = "radix order does not factor length " + std::to_string(n);
    return false;
  }
  // threads per sequence: every pass must fit kMaxE elements per thread (a
  // direct-DFT pass of a prime above 16 holds its n / tpt outputs, at most 8)
  constexpr int kMaxE = 16;
  int tpt = std::max(1, n / std::min(rmax, kMaxE));
  for (;; tpt++) {
    bool ok = true;
    for (int r : radices) {
      int nr = n / r;
      int per = (nr + tpt - 1) / tpt;
      if (r > kMaxE ? (n + tpt - 1) / tpt > 8 : per * r > kMaxE) ok = false;  // 8: kMaxDft
    }
    if (ok) break;
  }
  if (tpt > 1024) {
    if (why) *why = "length " + std::to_string(n) + " needs more than 1024 threads per sequence";
    return false;
  }
  // Wave-local passes (seq_threads): the threads of a sequence fit one wave
  // when tpt is a power of two up to 64; tpt is rounded up to one where that
  // idles at most a quarter of the lanes.  The caller decides whether the
  // stage may use it (LDS strides of its sequences).
  d.wave = 0;
  if (want_wave) {
    int t2 = 1;
    while (t2 < tpt) t2 *= 2;
    if (t2 <= 64 && 4 * tpt >= 3 * t2) {
      tpt = t2;
      d.wave = 1;
    }
  }
  d.tpt = tpt;
  d.npass = (int)radices.size();
  int ns = 1;
  for (int i = 0; i < d.npass; i++) {
    PassDesc& p = d.pass[i];
    p.r = radices[i];
    p.ns = ns;
    p.nr = n / p.r;
    p.twm = n / (ns * p.r);
    p.nsdiv = make_fastdiv((uint32_t)ns);
    ns *= p.r;
  }
  return true;
}

Plan::~Plan() {
  for (int d = 0; d < 3; d++) {
    if (tw[d]) (void)hipFree(tw[d]);
    if (coef[d]) (void)hipFree(coef[d]);
    if (scoef[d]) (void)hipFree(scoef[d]);
  }
}

template <typename T>
static Status upload_tables(Plan& pl) {
  const long double PI = 3.141592653589793238462643383279502884L;
  for (int d = 0; d < pl.dim; d++) {
    const int n = pl.shape[d];
    std::vector<T> tw(2 * (size_t)n);
    for (int j = 0; j < n; j++) {
      long double a = 2.0L * PI * (long double)j / (long double)n;
      tw[2 * j] = (T)cosl(a);
      tw[2 * j + 1] = (T)(-sinl(a));
    }
    // phi = 2(1-cos b)/h^2 = 4 sin^2(b/2)/h^2, chi = (2+cos b)/3,
    // psi = sin b / h with b = 2 pi k / N  (bri17.hpp:258-264)
    std::vector<T> cf(4 * (size_t)n);
    const long double h = (long double)pl.L[d] / (long double)n;
    for (int k = 0; k < n; k++) {
      long double sh = sinl(PI * (long double)k / (long double)n);
      long double b = 2.0L * PI * (long double)k / (long double)n;
      cf[4 * k + 0] = (T)(4.0L * sh * sh / (h * h));
      cf[4 * k + 1] = (T)(1.0L - 2.0L * sh * sh / 3.0L);
      cf[4 * k + 2] = (T)(sinl(b) / h);
      cf[4 * k + 3] = (T)0;
    }
    // strain-displacement factors c = cos a, sn = sin a, sh = sin a / h with
    // a = pi k / N (bri17.hpp:216-224)
    std::vector<T> sf(4 * (size_t)n);
    for (int k = 0; k < n; k++) {
      const long double a = PI * (long double)k / (long double)n;
      sf[4 * k + 0] = (T)cosl(a);
      sf[4 * k + 1] = (T)sinl(a);
      sf[4 * k + 2] = (T)(sinl(a) / h);
      sf[4 * k + 3] = (T)0;
    }
    hipError_t e = hipMalloc(&pl.scoef[d], sf.size() * sizeof(T));
    if (e != hipSuccess) return fail(3, std::string("hipMalloc(strain table): ") + hipGetErrorString(e));
    e = hipMemcpy(pl.scoef[d], sf.data(), sf.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(3, std::string("hipMemcpy(strain table): ") + hipGetErrorString(e));
    e = hipMalloc(&pl.tw[d], tw.size() * sizeof(T));
    if (e != hipSuccess) return fail(3, std::string("hipMalloc(twiddles): ") + hipGetErrorString(e));
    e = hipMalloc(&pl.coef[d], cf.size() * sizeof(T));
    if (e != hipSuccess) return fail(3, std::string("hipMalloc(modal table): ") + hipGetErrorString(e));
    e = hipMemcpy(pl.tw[d], tw.data(), tw.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(pl.coef[d], cf.data(), cf.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(3, std::string("hipMemcpy(tables): ") + hipGetErrorString(e));
  }
  return Status{};
}

static int round_up(int x, int m) { return (x + m - 1) / m * m; }

static const std::vector<int>& pick(const Tuning& t, int role, const std::vector<int>& dflt) {
  return t.radices[role].empty() ? dflt : t.radices[role];
}

Status build_plan(int dim, const int* shape, const double* L, int dtype, const Tuning& tun_in,
                  std::shared_ptr<Plan>& out) {
  Tuning tun = tun_in;
  StaticGeom sg;
  const bool use_static = dim == 3 && !tun.force_generic && tun.radices[0].empty() &&
                          tun.radices[1].empty() && tun.radices[2].empty() &&
                          static_supported(dtype, shape[0], shape[1], shape[2], &sg);
  if (use_static) {
    // the generic description has the split and tile width of the fixed-size path
    tun.plane_split = sg.p;
    tun.pencil_width = sg.w;
  }
  if (dim != 2 && dim != 3) return fail(1, "dim must be 2 or 3");
  if (dtype != 0 && dtype != 1) return fail(1, "dtype must be BRI17_F32 or BRI17_F64");
  auto pl = std::make_shared<Plan>();
  pl->dim = dim;
  pl->dtype = dtype;
  pl->esz = dtype == 0 ? 4 : 8;
  pl->tuning = tun;
  int64_t nn = 1;
  double hv = 1.0;
  for (int d = 0; d < dim; d++) {
    if (shape[d] < 1) return fail(1, "grid shape must be positive");
    if (!(L[d] > 0.0) || !std::isfinite(L[d])) return fail(1, "grid lengths must be positive and finite");
    pl->shape[d] = shape[d];
    pl->L[d] = L[d];
    nn *= shape[d];
    hv *= L[d] / shape[d];
  }
  if ((int64_t)dim * nn >= (int64_t(1) << 31))
    return fail(1, "DIM*|N| must be < 2^31 (the reference's int sizes, bri17.hpp:48)");
  pl->nnodes = nn;
  pl->cell_volume = hv;
  const size_t csz = 2 * pl->esz;  // complex element
  const size_t budget = std::min<size_t>(tun.lds_budget, 160 * 1024);

  const int nl = shape[dim - 1];
  const int even = nl % 2 == 0;
  const int m = even ? nl / 2 : nl;
  const int nlh = nl / 2 + 1;
  const int rsl = even ? m + 1 : nl;
  std::string why;

  PlaneParams& pp = pl->plane;
  std::memset(&pp, 0, sizeof(pp));
  pp.dim = dim;
  pp.g3 = 1;
  pp.nl = nl;
  pp.nlh = nlh;
  pp.even = even;
  pp.m = m;
  pp.rsl = rsl;
  pp.rs = nlh;
  if (!make_fft_desc(m, nl / m, pick(tun, 0, default_radices(m)), pp.row, &why, true))
    return fail(2, "row FFT: " + why);
  pp.mdiv = make_fastdiv((uint32_t)m);
  pp.hdiv = make_fastdiv((uint32_t)(m / 2 + 1));
  pp.nlhdiv = make_fastdiv((uint32_t)nlh);

  PencilParams& pc = pl->pencil;
  std::memset(&pc, 0, sizeof(pc));
  pc.dim = dim;
  pc.n0 = shape[0];
  pc.rs = nlh;
  pc.nlh = nlh;
  pc.xcd_remap = tun.xcd_remap >= 0 ? tun.xcd_remap : (use_static ? 0 : 1);

  if (dim == 3) {
    const int n1 = shape[1];
    // Smallest split P of the middle axis whose Q = N1/P rows fit in LDS.
    // In order of preference: the plane within kPlaneLds (several plane
    // workgroups share a CU and the batch kernel's items stay small) with
    // pencil tiles of at least 8 columns (64-byte rows of DIM N0 P entries);
    // any plane with such tiles; then narrower tiles, down to one column.
    // Any divisor up to 64 is a candidate; splits without a fused butterfly
    // run a Stockham P-point transform in the pencil.  Both kernels keep
    // their twiddle tables in LDS after their data.
    int P = 0;
    const size_t plane_tab = (size_t)(nl + n1) * csz;
    auto pencil_tab = [&](int c) { return (size_t)(shape[0] + (c > 1 && !fused_split(c) ? n1 : 0)) * csz; };
    std::vector<int> cands = {1};
    for (int c = 2; c <= 64 && c <= n1; c++)
      if (n1 % c == 0 && !default_radices(c).empty()) cands.push_back(c);
    const size_t pencil_col = (size_t)dim * shape[0] * csz;  // one column, per unit of P
    const struct { size_t plane; int min_w; } prefs[] = {
        {std::min(budget, kPlaneLds), 8}, {budget, 8}, {std::min(budget, kPlaneLds), 1}, {budget, 1}};
    for (const auto& pr : prefs) {
      for (int c : cands) {
        if (tun.plane_split && c != tun.plane_split) continue;
        if ((size_t)(n1 / c) * rsl * csz + plane_tab <= pr.plane &&
            pencil_col * c * pr.min_w + pencil_tab(c) <= budget) {
          P = c;
          break;
        }
      }
      if (P) break;
    }
    if (!P) return fail(2, "no middle-axis split fits the LDS budget for this grid");
    const int Q = n1 / P;
    pp.q = Q;
    pp.p = P;
    pp.n1 = n1;
    pp.has_col = 1;
    // wave-local column passes step through a column at the row stride
    // rsl: conflict-free LDS banks need it odd
    if (!make_fft_desc(Q, P, pick(tun, 1, default_radices(Q)), pp.col, &why, rsl % 2 == 1))
      return fail(2, "middle-axis FFT: " + why);
    int nthr = std::max(Q * pp.row.tpt, nlh * pp.col.tpt);
    nthr = std::min(1024, round_up(nthr, 64));
    nthr = std::max(nthr, round_up(std::max(pp.row.tpt, pp.col.tpt), 64));
    pl->plane_threads = nthr;
    pl->plane_blocks = dim * shape[0] * P;
    pl->plane_lds = (size_t)Q * rsl * csz + plane_tab;
    pc.p = P;
    pc.q = Q;
    pc.f = Q * nlh;
    if (!fused_split(P) && !make_fft_desc(P, Q, default_radices(P), pc.pfft, &why))
      return fail(2, "split transform: " + why);
  } else {
    const int nrows = dim * shape[0];
    int G = std::max(1, std::min(256 / pp.row.tpt, (nrows + 255) / 256));
    const size_t plane_tab = (size_t)nl * csz;
    if (rsl * csz + plane_tab > budget) return fail(2, "a row does not fit the LDS budget for this grid");
    G = std::min<int>(G, (int)((budget - plane_tab) / (rsl * csz)));
    G = std::max(1, std::min(G, nrows));
    pp.q = G;
    pp.p = 1;
    pp.n1 = G;
    pp.has_col = 0;
    pp.nrows = nrows;
    std::memset(&pp.col, 0, sizeof(pp.col));
    int nthr = std::min(1024, round_up(G * pp.row.tpt, 64));
    pl->plane_threads = nthr;
    pl->plane_blocks = (nrows + G - 1) / G;
    pl->plane_lds = (size_t)G * rsl * csz + plane_tab;
    pc.p = 1;
    pc.q = 1;
    pc.f = nlh;
  }

  // pencil tile width: widest power of two that fits the LDS budget
  if (!make_fft_desc(shape[0], 1, pick(tun, 2, default_radices(shape[0])), pc.ax0, &why))
    return fail(2, "axis-0 FFT: " + why);
  const size_t per_w = (size_t)dim * shape[0] * pc.p * csz;
  const size_t ptab = (size_t)pencil_table_entries(pc) * csz;
  int W = tun.pencil_width;
  if (!W) {
    W = 1;
    while (W < 64 && per_w * (2 * W) + ptab <= budget && 2 * W <= round_up(pc.f, 1)) W *= 2;
  }
  if (per_w * W + ptab > budget) return fail(2, "pencil tile does not fit the LDS budget for this grid");
  pc.w = W;
  pc.ntiles = (pc.f + W - 1) / W;
  pc.wdiv = make_fastdiv((uint32_t)W);
  pc.pwdiv = make_fastdiv((uint32_t)(pc.p * W));
  pc.rsdiv = make_fastdiv((uint32_t)nlh);
  {
    int nseq = dim * pc.p * W;
    int nthr = std::min(1024, round_up(nseq * pc.ax0.tpt, 64));
    nthr = std::max(nthr, round_up(pc.ax0.tpt, 64));
    if (!fused_split(pc.p)) nthr = std::max(nthr, round_up(pc.pfft.tpt, 64));
    pl->pencil_threads = nthr;
    pl->pencil_lds = per_w * W + ptab;
  }
  pl->ws_bytes = (size_t)dim * (nn / nl) * nlh * csz;
  pl->use_static = use_static;

  Status s = dtype == 0 ? upload_tables<float>(*pl) : upload_tables<double>(*pl);
  if (!s.ok()) return s;
  pp.tw_row = pl->tw[dim - 1];
  pp.tw_col = dim == 3 ? pl->tw[1] : nullptr;
  pc.tw0 = pl->tw[0];
  pc.tw1 = dim == 3 ? pl->tw[1] : nullptr;
  for (int d = 0; d < 3; d++) pc.coef[d] = d < dim ? pl->coef[d] : nullptr;
  for (int d = 0; d < 3; d++) pc.scoef[d] = d < dim ? pl->scoef[d] : nullptr;
  out = pl;
  return Status{};
}

void batch_minimum(const Plan& pl, size_t& lds, int& threads) {
  const size_t csz = 2 * pl.esz;
  const PlaneParams& pp = pl.plane;
  const PencilParams& pc = pl.pencil;
  const size_t plane = (pp.has_col ? (size_t)pp.q * pp.rsl * csz : (size_t)pp.rsl * csz) +
                       (size_t)plane_table_entries(pp) * csz;
  const size_t pencil = (size_t)pl.dim * pc.n0 * pc.p * csz + (size_t)pencil_table_entries(pc) * csz;
  lds = std::max(plane, pencil);
  int t = std::max({pp.row.tpt, pp.has_col ? pp.col.tpt : 1, pc.ax0.tpt,
                    fused_split(pc.p) ? 1 : pc.pfft.tpt});
  threads = round_up(t, 64);
}

bool batch_geometry(const Plan& pl, size_t lds, int threads, PlaneParams& pp, int& plane_blocks,
                    PencilParams& pc) {
  const size_t csz = 2 * pl.esz;
  pp = pl.plane;
  plane_blocks = pl.plane_blocks;
  const size_t plane_tab = (size_t)plane_table_entries(pp) * csz;
  if (plane_tab >= lds) return false;
  if (!pp.has_col) {
    // 2D: as many rows per item as fit the LDS and the workgroup transforms
    // in one round.  Odd rows are transformed in pairs (2j, 2j+1) counted
    // from the item's first row, so the row group stays even -- the pairs of
    // the plan's own even row group -- and a plan with an odd row group keeps
    // it.
    const int fit = (int)((lds - plane_tab) / ((size_t)pp.rsl * csz));
    int G = std::min(fit, pp.q);
    if (pp.even || pp.q % 2 == 0) {
      const int per_row = pp.even ? pp.row.tpt : (pp.row.tpt + 1) / 2;
      G = std::min({fit, pp.nrows, std::max(pp.q, threads / std::max(1, per_row))});
      if (!pp.even && G > 1) G &= ~1;
    }
    if (G < 1) return false;
    pp.q = G;
    pp.n1 = G;
    plane_blocks = (pp.nrows + G - 1) / G;
  } else if ((size_t)pp.q * pp.rsl * csz + plane_tab > lds) {
    return false;
  } else if (pp.p == 1 && (pp.even || pp.q % 2 == 0)) {
    // Several consecutive planes per item while the workgroup has threads
    // for all their row and column sequences in one round and the LDS holds
    // them.  (Odd rows are transformed in pairs: an item must not pair the
    // last row of one plane with the first of the next, hence even Q.)
    const int planes = pl.plane_blocks;  // components * N0 with split 1
    const int rseq = pp.even ? pp.q : pp.q / 2;
    const int per = std::max(rseq * pp.row.tpt, pp.nlh * pp.col.tpt);
    int g = std::min<int>(planes, (int)((lds - plane_tab) / ((size_t)pp.q * pp.rsl * csz)));
    g = std::min(g, std::max(1, threads / per));
    if (g > 1) {
      pp.g3 = g;
      pp.nplanes = planes;
      plane_blocks = (planes + g - 1) / g;
    }
  }
  const size_t per_w = (size_t)pl.dim * pc.n0 * pc.p * csz;
  const size_t ptab = (size_t)pencil_table_entries(pc) * csz;
  // The widest tile that fits, whatever the plan's own width: a column's
  // arithmetic does not depend on the tile it is in, and the batch's
  // workgroups have threads for wide tiles.
  int W = 1;
  while (W < 64 && W < pc.f) W *= 2;
  while (W > 1 && per_w * W + ptab > lds) W /= 2;
  if (per_w * W + ptab > lds) return false;
  pc.w = W;
  pc.ntiles = (pc.f + W - 1) / W;
  pc.wdiv = make_fastdiv((uint32_t)W);
  pc.pwdiv = make_fastdiv((uint32_t)(pc.p * W));
  return true;
}

static std::string radices_json(const FftDesc& d) {
  std::ostringstream o;
  o << "[";
  for (int i = 0; i < d.npass; i++) o << (i ? "," : "") << d.pass[i].r;
  o << "]";
  return o.str();
}

std::string describe_plan(const Plan& pl) {
  std::ostringstream o;
  const PlaneParams& pp = pl.plane;
  const PencilParams& pc = pl.pencil;
  o << "{\"dim\":" << pl.dim << ",\"fixed_size_kernels\":" << (pl.use_static ? "true" : "false")
    << ",\"dtype\":\"" << (pl.dtype == 0 ? "f32" : "f64") << "\""
    << ",\"shape\":[" << pl.shape[0] << "," << pl.shape[1];
  if (pl.dim == 3) o << "," << pl.shape[2];
  o << "],\"plane\":{\"split\":" << pp.p << ",\"rows\":" << pp.q << ",\"blocks\":" << pl.plane_blocks
    << ",\"threads\":" << pl.plane_threads << ",\"lds\":" << pl.plane_lds
    << ",\"row_fft\":{\"n\":" << pp.row.n << ",\"tpt\":" << pp.row.tpt
    << ",\"radices\":" << radices_json(pp.row) << "}";
  if (pp.has_col)
    o << ",\"mid_fft\":{\"n\":" << pp.col.n << ",\"tpt\":" << pp.col.tpt
      << ",\"radices\":" << radices_json(pp.col) << "}";
  o << "},\"pencil\":{\"width\":" << pc.w << ",\"tiles\":" << pc.ntiles
    << ",\"threads\":" << pl.pencil_threads << ",\"lds\":" << pl.pencil_lds
    << ",\"xcd_remap\":" << pc.xcd_remap << ",\"axis0_fft\":{\"n\":" << pc.ax0.n
    << ",\"tpt\":" << pc.ax0.tpt << ",\"radices\":" << radices_json(pc.ax0) << "}}"
    << ",\"workspace_bytes\":" << pl.ws_bytes;
  if (pl.use_static)
    o << ",\"fixed_size\":"
      << static_describe(pl.dtype, pl.shape[0], pl.shape[1], pl.shape[2], pl.tuning.plane_prefetch);
  o << "}";
  return o.str();
}

PencilParams pencil_for(const Plan& pl, double mu, double nu, int mode) {
  PencilParams pc = pl.pencil;
  pc.mode = mode;
  const double N = (double)pl.nnodes;
  // apply: K u scaled by |h|/|N| (tests/test_bri17.cpp:98-106); solve: the
  // inverse scaled by 1/(|h||N|).  The factor is folded into mu and
  // s = mu/(1-2nu) (bri17.hpp:266): c K(mu, s) = K(c mu, c s).
  pc.scale = mode == 0 ? pl.cell_volume / N : 1.0 / (pl.cell_volume * N);
  const double s = mu / (1.0 - 2.0 * nu);
  pc.mu = mode == 0 ? mu * pc.scale : mu / pc.scale;
  pc.s = mode == 0 ? s * pc.scale : s / pc.scale;
  pc.kinv = 1.0;
  if (mode >= 2) {
    // strain / eigenstress: the per-k map of the reference (K^ itself,
    // bri17.hpp:247-292), inverse DFT divided by |N| (tests/test_bri17.cpp:230-237)
    pc.scale = 1.0 / N;
    pc.mu = mu;
    pc.s = s;
    pc.bscale = 1.0 / N;
  }
  if (mode == 1 || mode == 3) {
    // The inverted block is built normalised: every entry of K(mu, s) is at
    // most (s + DIM mu) max_d phi_d <= (s + DIM mu) 4 / h_min^2, so dividing
    // mu and s by k = (s + mu) 4 / h_min^2 leaves entries of order one and a
    // determinant well inside the fp32 range whatever the units of mu and L.
    // The inverse is then scaled back by kinv = 1 / k (one product per
    // frequency, on the reciprocal of the determinant).
    double hmin2 = 0.0;
    for (int d = 0; d < pl.dim; d++) {
      const double h = pl.L[d] / pl.shape[d];
      hmin2 = d == 0 ? h * h : std::min(hmin2, h * h);
    }
    const double k = (std::fabs(pc.mu) + std::fabs(pc.s)) * 4.0 / hmin2;
    if (k > 0.0 && std::isfinite(k)) {
      pc.mu /= k;
      pc.s /= k;
      pc.kinv = 1.0 / k;
    }
  }
  return pc;
}

Status mode_geometry(const Plan& pl, double mu, double nu, int mode, ModeGeometry& g) {
  const int cin = mode_cin(pl.dim, mode), cout = mode_cout(pl.dim, mode);
  const int cmax = std::max(cin, cout);
  const size_t csz = 2 * pl.esz;
  auto plane = [&](int ncomp, PlaneParams& pp, int& blocks) {
    pp = pl.plane;
    if (pp.has_col) {
      blocks = ncomp * pl.shape[0] * pp.p;
    } else {
      pp.nrows = ncomp * pl.shape[0];
      blocks = (pp.nrows + pp.q - 1) / pp.q;
    }
  };
  plane(cin, g.fwd, g.fwd_blocks);
  plane(cout, g.inv, g.inv_blocks);
  g.plane_threads = pl.plane_threads;
  g.plane_lds = pl.plane_lds;
  g.pc = pencil_for(pl, mu, nu, mode);
  g.pencil_threads = pl.pencil_threads;
  g.pencil_lds = pl.pencil_lds;
  g.ws_bytes = pl.ws_bytes;
  if (mode <= 1) return Status{};
  PencilParams& pc = g.pc;
  const size_t per_w = (size_t)cmax * pc.n0 * pc.p * csz;
  const size_t ptab = (size_t)pencil_table_entries(pc) * csz;
  const size_t budget = std::min<size_t>(pl.tuning.lds_budget, 160 * 1024);
  int W = pc.w;
  while (W > 1 && per_w * W + ptab > budget) W /= 2;
  if (per_w * W + ptab > budget)
    return fail(2, "the pencil tile of this grid does not fit the LDS budget with " +
                       std::to_string(cmax) + " components");
  pc.w = W;
  pc.ntiles = (pc.f + W - 1) / W;
  pc.wdiv = make_fastdiv((uint32_t)W);
  pc.pwdiv = make_fastdiv((uint32_t)(pc.p * W));
  int nthr = std::min(1024, round_up(cmax * pc.p * W * pc.ax0.tpt, 64));
  nthr = std::max(nthr, round_up(pc.ax0.tpt, 64));
  if (!fused_split(pc.p)) nthr = std::max(nthr, round_up(pc.pfft.tpt, 64));
  g.pencil_threads = nthr;
  g.pencil_lds = per_w * W + ptab;
  g.ws_bytes = pl.ws_bytes / pl.dim * cmax;
  return Status{};
}

// Modes 2 and 3: runtime-size kernels over the mode's component counts.
static Status run_mode(const Plan& pl, double mu, double nu, int mode, const void* in, void* out,
                       void* ws, hipStream_t st, hipEvent_t* ev) {
  ModeGeometry g;
  Status s = mode_geometry(pl, mu, nu, mode, g);
  if (!s.ok()) return s;
  hipError_t e = hipSuccess;
  auto mark = [&](int i) {
    if (ev && e == hipSuccess) e = hipEventRecord(ev[i], st);
  };
  mark(0);
  if (pl.dtype == 0) {
    e = launch_plane_fwd<float>((const float*)in, ws, g.fwd, g.fwd_blocks, g.plane_threads,
                                g.plane_lds, st);
    mark(1);
    if (e == hipSuccess) e = launch_pencil<float>(ws, g.pc, g.pencil_threads, g.pencil_lds, st);
    mark(2);
    if (e == hipSuccess)
      e = launch_plane_inv<float>(ws, (float*)out, g.inv, g.inv_blocks, g.plane_threads,
                                  g.plane_lds, st);
  } else {
    e = launch_plane_fwd<double>((const double*)in, ws, g.fwd, g.fwd_blocks, g.plane_threads,
                                 g.plane_lds, st);
    mark(1);
    if (e == hipSuccess) e = launch_pencil<double>(ws, g.pc, g.pencil_threads, g.pencil_lds, st);
    mark(2);
    if (e == hipSuccess)
      e = launch_plane_inv<double>(ws, (double*)out, g.inv, g.inv_blocks, g.plane_threads,
                                   g.plane_lds, st);
  }
  mark(3);
  if (e != hipSuccess) return fail(3, std::string("kernel launch: ") + hipGetErrorString(e));
  return Status{};
}

Status run_pipeline(const Plan& pl, double mu, double nu, int mode, const void* in, void* out,
                    void* ws, hipStream_t st, hipEvent_t* ev) {
  if (mode >= 2) return run_mode(pl, mu, nu, mode, in, out, ws, st, ev);
  const PencilParams pc = pencil_for(pl, mu, nu, mode);
  hipError_t e = hipSuccess;
  if (pl.use_static) {
    e = launch_static(pl.dtype, pl.shape[0], pl.shape[1], pl.shape[2], mode, in, out, ws, pc,
                      pl.plane.tw_row, pl.plane.tw_col, pl.tuning.plane_prefetch, st, ev);
    if (e != hipSuccess) return fail(3, std::string("kernel launch: ") + hipGetErrorString(e));
    return Status{};
  }
  auto mark = [&](int i) {
    if (ev && e == hipSuccess) e = hipEventRecord(ev[i], st);
  };
  mark(0);
  if (pl.dtype == 0) {
    if (e == hipSuccess)
      e = launch_plane_fwd<float>((const float*)in, ws, pl.plane, pl.plane_blocks,
                                  pl.plane_threads, pl.plane_lds, st);
    mark(1);
    if (e == hipSuccess) e = launch_pencil<float>(ws, pc, pl.pencil_threads, pl.pencil_lds, st);
    mark(2);
    if (e == hipSuccess)
      e = launch_plane_inv<float>(ws, (float*)out, pl.plane, pl.plane_blocks, pl.plane_threads,
                                  pl.plane_lds, st);
  } else {
    if (e == hipSuccess)
      e = launch_plane_fwd<double>((const double*)in, ws, pl.plane, pl.plane_blocks,
                                   pl.plane_threads, pl.plane_lds, st);
    mark(1);
    if (e == hipSuccess) e = launch_pencil<double>(ws, pc, pl.pencil_threads, pl.pencil_lds, st);
    mark(2);
    if (e == hipSuccess)
      e = launch_plane_inv<double>(ws, (double*)out, pl.plane, pl.plane_blocks,
                                   pl.plane_threads, pl.plane_lds, st);
  }
  mark(3);
  if (e != hipSuccess) return fail(3, std::string("kernel launch: ") + hipGetErrorString(e));
  return Status{};
}

}  // namespace b17
