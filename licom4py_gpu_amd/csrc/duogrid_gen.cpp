// Host-side generator of the equal-edge gnomonic cubed-sphere duo-grid tables
// that the barotropic step consumes (the 14 "essential" fields).
//
// Restates the reference's Fortran grid library (src/duogrid_c/chtholly/...)
// for the fields listed in src/initial_field/main.py:30-39, evaluated on the
// a-grid window isd=jsd=-2 .. nx+3 (initial_field/main.py:54-55) and laid out
// C-contiguous as the reference's .npz after its transposes
// (initial_field/main.py:57-79, 116-124):
//   2-D fields  (6, N, N)          N = nx + 6, [t][p][q], p = i + 2, q = j + 2
//   a_pt        (6, N, N, 2)       (lon, lat)
//   k2e_coef    (6, N, N, 2)
//   a_gct/a_c2l/a_l2c (6, N, N, 2, 2)   [..][a][b] = Fortran (a+1, b+1)
//   k2e_loc_i/j (6, N, N) int32    = k2e_loc - isd
//
// Only libm (glibc) and IEEE +,-,*,/,sqrt are used, with the reference's
// operation order; compile with -ffp-contract=off so nothing is fused.
//
// Structure (own design): the reference fills full super-grid arrays by
// symmetric scatter loops; here the same values are produced by evaluating
// each needed super-grid point on demand through index folds that are the
// inverses of those scatters.  The one-dimensional coordinate lines and the
// k2e tables are shared by all six faces and all four edges, exactly as the
// reference builds them from face 1's west halo.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

constexpr double kRadius = 6.3712e6;              // lib_grid.F90:16
constexpr double kPi = 3.14159265358979323846;    // lib_grid.F90:17
constexpr double kOmega = 7.292e-5;               // lib_grid.F90:19
constexpr int kNg = 3;                            // a-grid halo (ng=3 passed at main.py:107)

struct Vec3 { double x, y, z; };

inline Vec3 lonlat_to_cart(double lon, double lat) {   // lib_grid.F90:80-95
  return {std::cos(lat) * std::cos(lon), std::cos(lat) * std::sin(lon), std::sin(lat)};
}

inline Vec3 cross(const Vec3& a, const Vec3& b) {      // lib_grid.F90:191-199
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// dot_product as a running sum from zero (gfortran inline expansion order).
inline double dot(const Vec3& a, const Vec3& b) {
  double s = 0.0;
  s = s + a.x * b.x;
  s = s + a.y * b.y;
  s = s + a.z * b.z;
  return s;
}

inline Vec3 normalize(const Vec3& r) {                 // lib_grid.F90:178-189
  double d = std::sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
  return {r.x / d, r.y / d, r.z / d};
}

// Unit tangent at p1 along the great circle towards p2 (lib_grid.F90:161-176).
inline Vec3 unit_towards(double lon1, double lat1, double lon2, double lat2) {
  Vec3 c1 = lonlat_to_cart(lon1, lat1);
  Vec3 c2 = lonlat_to_cart(lon2, lat2);
  Vec3 p12 = cross(c1, c2);
  return normalize(cross(p12, c1));
}

// Spherical quadrilateral area from the four interior angles (lib_grid.F90:123-159).
double quad_area(const double* p1, const double* p2, const double* p3, const double* p4) {
  Vec3 c1 = lonlat_to_cart(p1[0], p1[1]);
  Vec3 c2 = lonlat_to_cart(p2[0], p2[1]);
  Vec3 c3 = lonlat_to_cart(p3[0], p3[1]);
  Vec3 c4 = lonlat_to_cart(p4[0], p4[1]);
  Vec3 n12 = normalize(cross(c1, c2));
  Vec3 n23 = normalize(cross(c2, c3));
  Vec3 n34 = normalize(cross(c3, c4));
  Vec3 n41 = normalize(cross(c4, c1));
  double a123 = std::acos(-dot(n12, n23));
  double a234 = std::acos(-dot(n23, n34));
  double a341 = std::acos(-dot(n34, n41));
  double a412 = std::acos(-dot(n41, n12));
  return (a123 + a234 + a341 + a412 - 2.0 * kPi) * (kRadius * kRadius);
}

class DuoGrid {
 public:
  explicit DuoGrid(int nx) : nx_(nx) {
    ie_ = 2 * nx + 1;                 // super-grid interior 1..2nx+1
    lo_ = 1 - 2 * kNg;                // super-grid halo is 2*ng wide
    hi_ = ie_ + 2 * kNg;
    nc_ = 1 + nx;
    m_ = hi_ - lo_ + 1;
    build_line();
  }

  int nx() const { return nx_; }

  // Face-local gnomonic point at super-grid index (a, b) of face t (1..6),
  // as (lon, lat) (gen_lonlat.F90:110-152 with lib_cart2lonlat).
  void point(int t, int a, int b, double* lon, double* lat) const {
    const double x = line(a), y = line(b);
    double c[3];
    switch (t) {
      case 1: c[0] = 1.0; c[1] = x;    c[2] = y;    break;
      case 2: c[0] = -x;  c[1] = 1.0;  c[2] = y;    break;
      case 3: c[0] = -x;  c[1] = -y;   c[2] = 1.0;  break;
      case 4: c[0] = -1.0; c[1] = -y;  c[2] = -x;   break;
      case 5: c[0] = y;   c[1] = -1.0; c[2] = -x;   break;
      default: c[0] = y;  c[1] = x;    c[2] = -1.0; break;
    }
    // lib_cart2lonlat (lib_grid.F90:55-78)
    const double r = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    const double p0 = c[0] / r, p1 = c[1] / r, p2 = c[2] / r;
    double lo;
    if (std::fabs(p0) + std::fabs(p1) < 1.0e-5) {
      lo = 0.0;
    } else {
      lo = std::atan2(p1, p0);
    }
    if (lo < 0.0) lo = 2.0 * kPi + lo;
    *lon = lo;
    *lat = std::asin(p2);
  }

  double lat1(int a, int b) const { double lo, la; point(1, a, b, &lo, &la); return la; }

  // Kinked-grid point of face 1's west halo layer k at super index s: the
  // neighbouring face 5 interior point it copies (global_grid_util.F90:324-352,
  // odd face, source = west neighbour).
  double kik_lat_west(int k, int s) const {
    double lo, la;
    point(5, ie_ - (s - 1), ie_ + (1 - k - 1), &lo, &la);
    return la;
  }

  // dy on face 1 before the symmetric scatter (gen_cell.F90:60-64).
  double dy(int i, int j) const { return (lat1(i, j + 1) - lat1(i, j)) * kRadius; }
  // Folds inverting the scatters of gen_cell.F90:67-92 and 172-187.
  int fold_i(int a) const { return a <= nc_ ? a : ie_ + 1 - a; }
  int fold_j(int b) const { return b <= nc_ - 1 ? b : ie_ - b; }
  int fold_a(int a) const { return a <= nc_ - 1 ? a : ie_ - a; }
  double ext_dy(int a, int b) const { return dy(fold_i(a), fold_j(b)); }
  double ext_dx(int a, int b) const { return dy(fold_i(b), fold_j(a)); }
  double ext_da(int a, int b) const {
    const int i = fold_a(a), j = fold_a(b);
    double p1[2], p2[2], p3[2], p4[2];
    point(1, i, j, &p1[0], &p1[1]);
    point(1, i + 1, j, &p2[0], &p2[1]);
    point(1, i + 1, j + 1, &p3[0], &p3[1]);
    point(1, i, j + 1, &p4[0], &p4[1]);
    return quad_area(p1, p2, p3, p4);
  }

  int lo() const { return lo_; }
  int hi() const { return hi_; }
  int nc() const { return nc_; }
  int ie() const { return ie_; }

  double line(int s) const { return line_[s - lo_]; }

 private:
  // One-dimensional gnomonic coordinate line, equal-edge variant
  // (gen_lonlat.F90:77-108).
  void build_line() {
    line_.assign(m_, -999.0);
    const double alpha = std::asin(1.0 / std::sqrt(3.0));
    const double dela = 2.0 * alpha / static_cast<double>(ie_ - 1);
    auto at = [&](int s) -> double& { return line_[s - lo_]; };
    at(1) = -1.0;
    at(nc_) = 0.0;
    for (int j = 2; j <= nc_ - 1; ++j) {
      const double ang = static_cast<double>(j - 1) * dela - alpha;
      at(j) = std::tan(ang) * std::sqrt(2.0);
    }
    for (int j = lo_; j <= 0; ++j) {
      const int jj = 2 - j;
      const double ang = -0.5 * kPi - std::atan(at(jj));
      at(j) = std::tan(ang);
    }
    for (int j = lo_; j <= nc_ - 1; ++j) at(ie_ - j + 1) = -at(j);
  }

  int nx_, ie_, lo_, hi_, nc_, m_;
  std::vector<double> line_;
};

// Per-layer k2e remap tables (gen_k2e.F90:28-172 with Lagrange order 2,
// lib_grid.F90:213-231).  Layer ii = 1..3; along-edge a-grid index m.
struct K2ETables {
  int nx;
  // index [ii-1][m + kNg - 1], m in [1-kNg+1 .. nx+kNg]  (store -2..nx+3)
  std::vector<int> loc;
  std::vector<double> c0, c1;
  double min_margin;  // smallest |x(k) - y| met by a search comparison
  int at(int ii, int m) const { return (ii - 1) * (nx + 2 * kNg) + (m + 2); }
};

K2ETables build_k2e(const DuoGrid& g) {
  const int nx = g.nx();
  const int nc = g.nc();
  const int ie = g.ie();
  K2ETables T;
  T.nx = nx;
  const int n_along = nx + 2 * kNg;
  T.loc.assign(kNg * n_along, -999);
  T.c0.assign(kNg * n_along, -999.0);
  T.c1.assign(kNg * n_along, -999.0);
  T.min_margin = INFINITY;

  std::vector<double> kline(ie + 1), eline(g.hi() - g.lo() + 1);
  std::vector<double> x(nx + 1);
  for (int ii = 1; ii <= kNg; ++ii) {
    const int k = 2 * ii - 1;  // super-grid layer of a-grid halo layer ii
    // kinked-grid coordinate line (gen_coords.F90:222-232)
    const double kc = g.kik_lat_west(k, nc);
    kline[nc] = 0.0;
    for (int j = 1; j <= nc - 1; ++j) {
      kline[j] = g.kik_lat_west(k, j) - kc;
      kline[ie - j + 1] = -kline[j];
    }
    // extended-grid coordinate line (gen_coords.F90:248-258)
    auto el = [&](int s) -> double& { return eline[s - g.lo()]; };
    const double ec = g.lat1(1 - k, nc);
    el(nc) = 0.0;
    for (int j = g.lo(); j <= nc - 1; ++j) {
      el(j) = g.lat1(1 - k, j) - ec;
      el(ie - j + 1) = -el(j);
    }
    for (int kk = 1; kk <= nx; ++kk) x[kk] = kline[2 * kk];
    for (int m = 2 - ii; m <= nx + ii - 1; ++m) {
      const double yy = el(2 * m);
      int klo = 1, khi = nx;
      while (khi - klo > 1) {
        const int mid = (khi + klo) / 2;
        T.min_margin = std::min(T.min_margin, std::fabs(x[mid] - yy));
        if (x[mid] > yy) khi = mid; else klo = mid;
      }
      klo = std::max(klo, 1);
      klo = std::min(klo, nx - 1);
      khi = klo + 1;
      const double x1 = x[klo], x2 = x[khi];
      const int idx = T.at(ii, m);
      T.loc[idx] = klo;
      T.c0[idx] = 1.0 * (yy - x2) / (x1 - x2);
      T.c1[idx] = 1.0 * (yy - x1) / (x2 - x1);
    }
  }
  return T;
}

}  // namespace

extern "C" {

// Fills the essential duo-grid fields for all six faces (see header comment
// for layouts).  Any output pointer may be null to skip that field.  Returns 0
// on success, -1 on invalid nx.  *min_margin (if non-null) receives the
// smallest gap between a searched coordinate and a table node met while
// building k2e_loc: a near-tie audit for the integer table.
int duogrid_generate(int nx, double* a_pt, double* a_sina, double* a_f,
                     double* rdx, double* rdy, double* rda, double* a_gct,
                     double* a_c2l, double* a_l2c, double* c_dy, double* d_dx,
                     double* k2e_coef, int32_t* k2e_loc_i, int32_t* k2e_loc_j,
                     double* min_margin) {
  if (nx < 4 || nx > 8192) return -1;
  const DuoGrid g(nx);
  const K2ETables T = build_k2e(g);
  if (min_margin) *min_margin = T.min_margin;
  const int N = nx + 2 * kNg;
  const int isd = 1 - kNg;  // -2
  const size_t face = static_cast<size_t>(N) * N;

  // Face-independent metric fields: 1/dx, 1/dy, 1/da and edge lengths are
  // copied from face 1 to all faces by the reference (gen_cell.F90:67-92).
  std::vector<double> s_rdx(face), s_rdy(face), s_rda(face), s_cdy(face), s_ddx(face);
#pragma omp parallel for schedule(static)
  for (int p = 0; p < N; ++p) {
    const int i = p + isd;
    for (int q = 0; q < N; ++q) {
      const int j = q + isd;
      const size_t o = static_cast<size_t>(p) * N + q;
      // global_grid_base.F90:129-138, 160-161, 174-175
      const double adx = g.ext_dx(2 * i - 1, 2 * j) + g.ext_dx(2 * i, 2 * j);
      const double ady = g.ext_dy(2 * i, 2 * j - 1) + g.ext_dy(2 * i, 2 * j);
      const double ada = g.ext_da(2 * i - 1, 2 * j - 1) + g.ext_da(2 * i - 1, 2 * j) +
                         g.ext_da(2 * i, 2 * j - 1) + g.ext_da(2 * i, 2 * j);
      s_rda[o] = 1.0 / ada;
      s_rdx[o] = 1.0 / adx;
      s_rdy[o] = 1.0 / ady;
      s_cdy[o] = g.ext_dy(2 * i - 1, 2 * j - 1) + g.ext_dy(2 * i - 1, 2 * j);
      s_ddx[o] = g.ext_dx(2 * i - 1, 2 * j - 1) + g.ext_dx(2 * i, 2 * j - 1);
    }
  }

  for (int t = 1; t <= 6; ++t) {
    const size_t fo = static_cast<size_t>(t - 1) * face;
    if (rdx) std::memcpy(rdx + fo, s_rdx.data(), face * sizeof(double));
    if (rdy) std::memcpy(rdy + fo, s_rdy.data(), face * sizeof(double));
    if (rda) std::memcpy(rda + fo, s_rda.data(), face * sizeof(double));
    if (c_dy) std::memcpy(c_dy + fo, s_cdy.data(), face * sizeof(double));
    if (d_dx) std::memcpy(d_dx + fo, s_ddx.data(), face * sizeof(double));

#pragma omp parallel for schedule(static)
    for (int p = 0; p < N; ++p) {
      const int a = 2 * (p + isd);
      for (int q = 0; q < N; ++q) {
        const int b = 2 * (q + isd);
        const size_t o = fo + static_cast<size_t>(p) * N + q;
        double lon, lat;
        g.point(t, a, b, &lon, &lat);
        if (a_pt) { a_pt[2 * o] = lon; a_pt[2 * o + 1] = lat; }

        // covariant basis (gen_vec.F90:53-82)
        double ln, lt;
        Vec3 e1, e2;
        if (a == g.hi()) {
          g.point(t, a - 1, b, &ln, &lt);
          Vec3 e = unit_towards(lon, lat, ln, lt);
          e1 = {-e.x, -e.y, -e.z};
        } else {
          g.point(t, a + 1, b, &ln, &lt);
          e1 = unit_towards(lon, lat, ln, lt);
        }
        if (b == g.hi()) {
          g.point(t, a, b - 1, &ln, &lt);
          Vec3 e = unit_towards(lon, lat, ln, lt);
          e2 = {-e.x, -e.y, -e.z};
        } else {
          g.point(t, a, b + 1, &ln, &lt);
          e2 = unit_towards(lon, lat, ln, lt);
        }
        // local east/north unit vectors (gen_vec.F90:269-276)
        const Vec3 elon = {-std::sin(lon), std::cos(lon), 0.0};
        const Vec3 elat = {-std::sin(lat) * std::cos(lon), -std::sin(lat) * std::sin(lon),
                           std::cos(lat)};
        // metric and rotation matrices (gen_vec.F90:338-341, 413-485)
        const double cosa = dot(e1, e2);
        const double sina = std::sqrt(1.0 - cosa * cosa);
        const double gco[2][2] = {{1.0, cosa}, {cosa, 1.0}};
        const double cf = 1.0 / sina / sina;
        const double gct[2][2] = {{cf, -(cf * cosa)}, {-(cf * cosa), cf}};
        const double m11 = dot(e1, elon), m12 = dot(e2, elon);
        const double m21 = dot(e1, elat), m22 = dot(e2, elat);
        const double ct2rll[2][2] = {{m11, m12}, {m21, m22}};
        const double cd = 1.0 / (m11 * m22 - m12 * m21);
        const double rll2ct[2][2] = {{cd * m22, -(cd * m12)}, {-(cd * m21), cd * m11}};
        if (a_sina) a_sina[o] = sina;
        for (int r = 0; r < 2; ++r) {
          for (int c = 0; c < 2; ++c) {
            // matmul accumulated from zero, k = 1 then 2
            double c2l = 0.0, l2c = 0.0;
            c2l = c2l + ct2rll[r][0] * gct[0][c];
            c2l = c2l + ct2rll[r][1] * gct[1][c];
            l2c = l2c + gco[r][0] * rll2ct[0][c];
            l2c = l2c + gco[r][1] * rll2ct[1][c];
            if (a_gct) a_gct[4 * o + 2 * r + c] = gct[r][c];
            if (a_c2l) a_c2l[4 * o + 2 * r + c] = c2l;
            if (a_l2c) a_l2c[4 * o + 2 * r + c] = l2c;
          }
        }
        // Coriolis with rotation angle alpha = 0 (global_grid_base.F90:193-202)
        if (a_f) {
          const double alpha = 0.0;
          a_f[o] = 2.0 * kOmega *
                   (-1.0 * std::cos(lon) * std::cos(lat) * std::sin(alpha) +
                    std::sin(lat) * std::cos(alpha));
        }
        if (k2e_coef) { k2e_coef[2 * o] = -999.0; k2e_coef[2 * o + 1] = -999.0; }
        if (k2e_loc_i) k2e_loc_i[o] = -999 - isd;
        if (k2e_loc_j) k2e_loc_j[o] = -999 - isd;
      }
    }

    // k2e tables on the four halo strips (gen_k2e.F90:310-326), identical on
    // every face and edge.
    for (int ii = 1; ii <= kNg; ++ii) {
      for (int m = 2 - ii; m <= nx + ii - 1; ++m) {
        const int idx = T.at(ii, m);
        const int cells[4][2] = {
            {m, 1 - ii}, {m, nx + ii}, {1 - ii, m}, {nx + ii, m}};
        for (const auto& c : cells) {
          const size_t o = fo + static_cast<size_t>(c[0] - isd) * N + (c[1] - isd);
          if (k2e_coef) { k2e_coef[2 * o] = T.c0[idx]; k2e_coef[2 * o + 1] = T.c1[idx]; }
          if (k2e_loc_i) k2e_loc_i[o] = T.loc[idx] - isd;
          if (k2e_loc_j) k2e_loc_j[o] = T.loc[idx] - isd;
        }
      }
    }
  }
  return 0;
}

// The one-dimensional super-grid coordinate line (for tests): writes
// line(s) for s = 1-2*ng .. 2*nx+1+2*ng into out (length 2*nx+13).
int duogrid_line(int nx, double* out) {
  if (nx < 4 || nx > 8192) return -1;
  const DuoGrid g(nx);
  for (int s = g.lo(); s <= g.hi(); ++s) out[s - g.lo()] = g.line(s);
  return 0;
}

}  // extern "C"
