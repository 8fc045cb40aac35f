/* FP64 CPU oracle of the Dioptrin integration hot path (test infrastructure
 * only; see oracle.h). Build: -O2 -ffp-contract=off -fno-fast-math. */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.14159265358979311599796346854418516159057617187500
#define OR_TAU 6.28318530717958623199592693708837032318115234375000
#define OR_EPS DBL_EPSILON

/* ---------------------------------------------------------------- helpers */

/* Rust `f64 as isize` / `as usize`: saturating, NaN -> 0. */
static long long sat_isize(double v) {
  if (v != v) return 0;
  if (v >= 9223372036854775807.0) return 9223372036854775807LL;
  if (v <= -9223372036854775808.0) return (-9223372036854775807LL - 1);
  return (long long)v;
}
static unsigned long long sat_usize(double v) {
  if (!(v > 0.0)) return 0; /* negative, zero or NaN */
  if (v >= 18446744073709551615.0) return 18446744073709551615ULL;
  return (unsigned long long)v;
}

/* Rust f64::rem_euclid: r = a % b; r < 0 ? r + |b| : r. */
static double rem_euclid(double a, double b) {
  double r = fmod(a, b);
  return r < 0.0 ? r + fabs(b) : r;
}

/* pyfai_1d.rs:180-186: x.rem_euclid(TAU), returning values already in
 * [0, TAU) unchanged. */
static double wrap_tau(double x) {
  if (x >= 0.0 && x < OR_TAU) return x;
  return rem_euclid(x, OR_TAU);
}

/* src/math/mod.rs:8-32: first element, then strict comparisons. */
static double vmin(const double *v, size_t n) {
  double m = v[0];
  for (size_t i = 1; i < n; i++)
    if (v[i] < m) m = v[i];
  return m;
}
static double vmax(const double *v, size_t n) {
  double m = v[0];
  for (size_t i = 1; i < n; i++)
    if (v[i] > m) m = v[i];
  return m;
}

/* Growable per-bin entry list. */
typedef struct {
  uint64_t *col;
  double *w;
  size_t len, cap;
} bucket;

static void bucket_push(bucket *b, uint64_t col, double w) {
  if (b->len == b->cap) {
    size_t cap = b->cap ? 2 * b->cap : 8;
    b->col = (uint64_t *)realloc(b->col, cap * sizeof(uint64_t));
    b->w = (double *)realloc(b->w, cap * sizeof(double));
    b->cap = cap;
  }
  b->col[b->len] = col;
  b->w[b->len] = w;
  b->len++;
}

/* pixel_split.rs:332-356: flatten per-bin lists to CSR in bin order. */
static void buckets_to_lut(bucket *bk, size_t n, int weighted, double min, double step,
                           or_lut *out) {
  out->nbins = n;
  out->row_ptr = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  out->row_ptr[0] = 0;
  for (size_t i = 0; i < n; i++) out->row_ptr[i + 1] = out->row_ptr[i] + bk[i].len;
  size_t nnz = out->row_ptr[n];
  out->nnz = nnz;
  out->col = (uint64_t *)malloc((nnz ? nnz : 1) * sizeof(uint64_t));
  out->w = weighted ? (double *)malloc((nnz ? nnz : 1) * sizeof(double)) : NULL;
  size_t k = 0;
  for (size_t i = 0; i < n; i++) {
    for (size_t e = 0; e < bk[i].len; e++, k++) {
      out->col[k] = bk[i].col[e];
      if (weighted) out->w[k] = bk[i].w[e];
    }
    free(bk[i].col);
    free(bk[i].w);
  }
  out->hist_min = min;
  out->hist_step = step;
}

void or_lut_free(or_lut *lut) {
  free(lut->row_ptr);
  free(lut->col);
  free(lut->w);
  memset(lut, 0, sizeof(*lut));
}

/* --------------------------------------------------------------- geometry */

typedef struct {
  double s1, c1, s2, c2, s3, c3;
} trig;

/* geometry.rs:640-660 (TrigonometricConstants::new). */
static trig make_trig(const or_poni *p) {
  trig t;
  t.s1 = sin(p->rot1);
  t.c1 = cos(p->rot1);
  t.s2 = sin(p->rot2);
  t.c2 = cos(p->rot2);
  t.s3 = sin(p->rot3);
  t.c3 = cos(p->rot3);
  return t;
}

/* geometry.rs:663-681, evaluated left to right exactly as written. */
static double f_x1(double p1, double p2, double p3, const trig *t) {
  return p1 * t->c2 * t->c3 + p2 * (t->c3 * t->s1 * t->s2 - t->c1 * t->s3) -
         p3 * (t->c1 * t->c3 * t->s2 + t->s1 * t->s3);
}
static double f_x2(double p1, double p2, double p3, const trig *t) {
  return p1 * t->c2 * t->s3 + p2 * (t->c1 * t->c3 + t->s1 * t->s2 * t->s3) -
         p3 * (-(t->c3 * t->s1) + t->c1 * t->s2 * t->s3);
}
static double f_x3(double p1, double p2, double p3, const trig *t) {
  return p1 * t->s2 - p2 * t->c2 * t->s1 + p3 * t->c1 * t->c2;
}
/* geometry.rs:684-692. */
static double f_tth(double x, double y, double z) { return atan2(sqrt(y * y + x * x), z); }
static double f_azimuth(double x, double y) { return atan2(x, y); }

/* geometry.rs:288-365 (det reference -> poni reference -> sample frame):
 * p1 = d1 * pixel1 - poni1, p2 = d2 * pixel2 - poni2, p3 = distance. */
static void point_xyz(const or_poni *p, const trig *t, double d1, double d2, double *x, double *y,
                      double *z) {
  double p1 = d1 * p->pixel1 - p->poni1;
  double p2 = d2 * p->pixel2 - p->poni2;
  *x = f_x1(p1, p2, p->distance, t);
  *y = f_x2(p1, p2, p->distance, t);
  *z = f_x3(p1, p2, p->distance, t);
}

void or_positions(const or_poni *p, const double *d1, const double *d2, size_t n, double *x,
                  double *y, double *z) {
  trig t = make_trig(p);
  for (size_t i = 0; i < n; i++) point_xyz(p, &t, d1[i], d2[i], &x[i], &y[i], &z[i]);
}

/* geometry.rs:519-549 (calc_tth_and_azimuth). */
void or_tth_chi(const or_poni *p, const double *d1, const double *d2, size_t n, double *tth,
                double *chi) {
  trig t = make_trig(p);
  for (size_t i = 0; i < n; i++) {
    double x, y, z;
    point_xyz(p, &t, d1[i], d2[i], &x, &y, &z);
    if (tth) tth[i] = f_tth(x, y, z);
    if (chi) chi[i] = f_azimuth(x, y);
  }
}

/* geometry.rs:229-248, 265-278: corner grid d1 = i, d2 = j. */
void or_corners(const or_poni *p, size_t dim1, size_t dim2, double *tth, double *chi) {
  trig t = make_trig(p);
  size_t k = 0;
  for (size_t i = 0; i <= dim1; i++)
    for (size_t j = 0; j <= dim2; j++, k++) {
      double x, y, z;
      point_xyz(p, &t, (double)i, (double)j, &x, &y, &z);
      tth[k] = f_tth(x, y, z);
      chi[k] = f_azimuth(x, y);
    }
}

/* geometry.rs:308-320, 445-487: centres d1 = i + 0.5, d2 = j + 0.5. */
void or_centres(const or_poni *p, size_t dim1, size_t dim2, double *tth, double *chi) {
  trig t = make_trig(p);
  size_t k = 0;
  for (size_t i = 0; i < dim1; i++)
    for (size_t j = 0; j < dim2; j++, k++) {
      double x, y, z;
      point_xyz(p, &t, (double)i + 0.5, (double)j + 0.5, &x, &y, &z);
      if (tth) tth[k] = f_tth(x, y, z);
      if (chi) chi[k] = f_azimuth(x, y);
    }
}

/* geometry.rs:552-617 (flat detector) and 697-712: cos(a)^3 with
 * cos(a) = p3 / sqrt(p1^2 + p2^2 + p3^2). */
void or_solid_angle(const or_poni *p, size_t dim1, size_t dim2, double *out) {
  size_t k = 0;
  for (size_t i = 0; i < dim1; i++)
    for (size_t j = 0; j < dim2; j++, k++) {
      double p1 = ((double)i + 0.5) * p->pixel1 - p->poni1;
      double p2 = ((double)j + 0.5) * p->pixel2 - p->poni2;
      double p3 = p->distance;
      double cosa = p3 / sqrt(p1 * p1 + p2 * p2 + p3 * p3);
      out[k] = cosa * cosa * cosa;
    }
}

/* geometry.rs:625-636, 715-717. */
void or_polarization(double factor, const double *tth, const double *chi, size_t n, double *out) {
  for (size_t i = 0; i < n; i++) {
    double c = cos(tth[i]);
    double c2 = c * c;
    out[i] = 0.5 * (1.0 + c2 - factor * cos(2.0 * chi[i]) * (1.0 - c2));
  }
}

/* corrections.rs:57-93: normalization = polarization * solid angle. */
void or_normalization(const or_poni *p, size_t dim1, size_t dim2, int has_pol, double factor,
                      double *out) {
  size_t n = dim1 * dim2;
  double *sa = (double *)malloc(n * sizeof(double));
  or_solid_angle(p, dim1, dim2, sa);
  if (has_pol) {
    double *tth = (double *)malloc(n * sizeof(double));
    double *chi = (double *)malloc(n * sizeof(double));
    or_centres(p, dim1, dim2, tth, chi);
    or_polarization(factor, tth, chi, n, out);
    for (size_t i = 0; i < n; i++) out[i] = out[i] * sa[i];
    free(tth);
    free(chi);
  } else {
    for (size_t i = 0; i < n; i++) out[i] = 1.0 * sa[i];
  }
  free(sa);
}

/* integration.rs:132-187 (RadialConverter::convert). Degrees use Rust's
 * f64::to_degrees, a product with the correctly rounded 180/pi. */
void or_convert_unit(int unit, double wavelength, const double *tth, size_t n, double *out) {
  double q_scale = 0.0;
  if (unit == OR_UNIT_Q_A || unit == OR_UNIT_Q_NM)
    q_scale = 4.0 * OR_PI / (wavelength * (unit == OR_UNIT_Q_A ? 1e10 : 1e9));
  for (size_t i = 0; i < n; i++) {
    switch (unit) {
      case OR_UNIT_TTH_RAD: out[i] = tth[i]; break;
      case OR_UNIT_TTH_DEG: out[i] = tth[i] * (180.0 / OR_PI); break;
      default: out[i] = q_scale * sin(tth[i] * 0.5); break;
    }
  }
}

void or_bins(size_t n, double min, double step, double *out) {
  for (size_t i = 0; i < n; i++) out[i] = min + ((double)i + 0.5) * step;
}

/* ---------------------------------------------------- pixel splitting, 1D */

/* pixel_split.rs:14-25: shoelace sum, then 0.5 *. */
static double polygon_area(const double (*poly)[2], int n) {
  if (n < 3) return 0.0;
  double area = 0.0;
  for (int i = 0; i < n; i++) {
    const double *a = poly[i];
    const double *b = poly[(i + 1) % n];
    area += a[0] * b[1] - b[0] * a[1];
  }
  return 0.5 * area;
}

/* pixel_split.rs:28-45. */
static void intersect_x(const double *p0, const double *p1, double xe, double *out) {
  double dx = p1[0] - p0[0];
  if (fabs(dx) < OR_EPS) {
    out[0] = xe;
    out[1] = p0[1];
    return;
  }
  double t = (xe - p0[0]) / dx;
  out[0] = xe;
  out[1] = p0[1] + t * (p1[1] - p0[1]);
}
static void intersect_y(const double *p0, const double *p1, double ye, double *out) {
  double dy = p1[1] - p0[1];
  if (fabs(dy) < OR_EPS) {
    out[0] = p0[0];
    out[1] = ye;
    return;
  }
  double t = (ye - p0[1]) / dy;
  out[0] = p0[0] + t * (p1[0] - p0[0]);
  out[1] = ye;
}

/* One Sutherland-Hodgman stage (pixel_split.rs:48-137). axis 0 = x, 1 = y;
 * keep_ge: inside means coordinate >= edge, else <= edge. */
static int clip_stage(const double (*in)[2], int n, int axis, int keep_ge, double edge,
                      double (*out)[2]) {
  if (n == 0) return 0;
  int m = 0;
  const double *prev = in[n - 1];
  int prev_in = keep_ge ? (prev[axis] >= edge) : (prev[axis] <= edge);
  for (int i = 0; i < n; i++) {
    const double *cur = in[i];
    int cur_in = keep_ge ? (cur[axis] >= edge) : (cur[axis] <= edge);
    if (cur_in) {
      if (!prev_in) {
        if (axis == 0) intersect_x(prev, cur, edge, out[m]);
        else intersect_y(prev, cur, edge, out[m]);
        m++;
      }
      out[m][0] = cur[0];
      out[m][1] = cur[1];
      m++;
    } else if (prev_in) {
      if (axis == 0) intersect_x(prev, cur, edge, out[m]);
      else intersect_y(prev, cur, edge, out[m]);
      m++;
    }
    prev = cur;
    prev_in = cur_in;
  }
  return m;
}

/* pixel_split.rs:140-163: left, right, bottom, top; stops at an empty
 * stage. Returns the vertex count of the clipped polygon in `res`. */
static int clip_rect(const double (*poly)[2], int n, double xmin, double xmax, double ymin,
                     double ymax, double (*res)[2]) {
  double a[24][2], b[24][2];
  int m = clip_stage(poly, n, 0, 1, xmin, a);
  if (m == 0) return 0;
  m = clip_stage((const double(*)[2])a, m, 0, 0, xmax, b);
  if (m == 0) return 0;
  m = clip_stage((const double(*)[2])b, m, 1, 1, ymin, a);
  if (m == 0) return 0;
  m = clip_stage((const double(*)[2])a, m, 1, 0, ymax, res);
  return m;
}

/* pixel_split.rs:166-175. */
static double unwrap_angle(double angle, double reference) {
  double a = angle;
  double delta = a - reference;
  if (delta > OR_PI) a -= 2.0 * OR_PI;
  else if (delta < -OR_PI) a += 2.0 * OR_PI;
  return a;
}

/* pyfai_1d.rs:188-201: beam pixel test on (radial, chi) corners. */
static int singular(const double (*poly)[2]) {
  double winding = 0.0;
  for (int i = 0; i < 4; i++) {
    if (poly[i][0] == 0.0) return 1;
    double delta = wrap_tau(poly[(i + 1) % 4][1] - poly[i][1] + OR_PI) - OR_PI;
    if (fabs(delta) >= OR_PI - 64.0 * OR_EPS) return 1;
    winding += delta;
  }
  return fabs(winding) > OR_PI;
}

/* integration/grid.rs:48-108 (IntegrationGrid2d::new) validity, for the
 * 1D sampling grid (n radial, 1 azimuthal, chi in (-pi, pi)). */
static int sample_grid_ok(size_t n, double lo, double hi) {
  double bounds[2][2] = {{lo, hi}, {-OR_PI, OR_PI}};
  size_t counts[2] = {n, 1};
  if (n == 0) return 0;
  for (int k = 0; k < 2; k++) {
    double l = bounds[k][0], h = bounds[k][1];
    double step = (h - l) / (double)counts[k];
    if (!isfinite(l) || !isfinite(h) || !isfinite(step) || step <= 0.0 || l + step == l ||
        h - step == h)
      return 0;
  }
  if (lo < 0.0) return 0;
  return 1;
}

/* pyfai_1d.rs:208-252: 128 x 128 x 4 detector-plane samples of one pixel,
 * binned on the grid (r0, r1) x (-pi, pi) with n radial bins. Appends
 * (pixel, hits / 65536) to the bins in ascending bin order. */
static void sample_pixel(const or_poni *p, const trig *t, size_t dim2, size_t pixel, int unit,
                         size_t n, double r0, double r1, int clamp_low, bucket *bk) {
  const int N = 128;
  const double off[4][2] = {{0.25, 0.5}, {0.5, 0.25}, {0.75, 0.5}, {0.5, 0.75}};
  const double count = (double)(N * N * 4);
  unsigned *hits = (unsigned *)calloc(n, sizeof(unsigned));
  double row0 = (double)(pixel / dim2), col0 = (double)(pixel % dim2);
  double a0 = -OR_PI, a1 = OR_PI;
  double q_scale = 0.0;
  if (unit == OR_UNIT_Q_A || unit == OR_UNIT_Q_NM)
    q_scale = 4.0 * OR_PI / (p->wavelength * (unit == OR_UNIT_Q_A ? 1e10 : 1e9));
  for (int row = 0; row < N; row++)
    for (int col = 0; col < N; col++)
      for (int o = 0; o < 4; o++) {
        double d1 = row0 + ((double)row + off[o][0]) / (double)N;
        double d2 = col0 + ((double)col + off[o][1]) / (double)N;
        double x, y, z;
        point_xyz(p, t, d1, d2, &x, &y, &z);
        double tth = f_tth(x, y, z);
        double a = f_azimuth(x, y);
        double r;
        switch (unit) {
          case OR_UNIT_TTH_RAD: r = tth; break;
          case OR_UNIT_TTH_DEG: r = tth * (180.0 / OR_PI); break;
          default: r = q_scale * sin(tth * 0.5); break;
        }
        a = a0 + rem_euclid(a - a0, OR_TAU);
        if ((r < r0 && !clamp_low) || r >= r1 || a >= a1) continue;
        unsigned long long ir = r < r0 ? 0 : sat_usize((r - r0) / (r1 - r0) * (double)n);
        unsigned long long ia = sat_usize((a - a0) / (a1 - a0) * 1.0);
        if (ir < n && ia < 1) hits[ia * n + ir]++;
      }
  for (size_t b = 0; b < n; b++)
    if (hits[b]) bucket_push(&bk[b], pixel, (double)hits[b] / count);
  free(hits);
}

/* pixel_split.rs:190-328 (lut1d_ps_fractions), serial and in pixel order,
 * which is the order the reference's fold/reduce assembles. */
static void ps_fractions(const or_poni *p, size_t dim1, size_t dim2, const double *radial_c,
                         const double *chi_c, int unit, size_t n, double min, double step,
                         const double *rc, const double *ac, int clamp_low, bucket *bk) {
  trig t = make_trig(p);
  double glo = min, ghi = min + (double)n * step;
  int grid_ok = sample_grid_ok(n, glo, ghi);
  size_t w = dim2 + 1;
  size_t npix = dim1 * dim2;
  for (size_t idx = 0; idx < npix; idx++) {
    size_t b = idx / dim2 * w + idx % dim2;
    size_t i00 = b, i01 = b + 1, i11 = b + w + 1, i10 = b + w;
    double r00 = rc[i00], r01 = rc[i01], r10 = rc[i10], r11 = rc[i11];
    if (grid_ok) {
      double raw[4][2] = {{rc[i00], ac[i00]}, {rc[i01], ac[i01]}, {rc[i11], ac[i11]},
                          {rc[i10], ac[i10]}};
      if (singular((const double(*)[2])raw)) {
        sample_pixel(p, &t, dim2, idx, unit, n, glo, ghi, clamp_low, bk);
        continue;
      }
    }
    double ca = chi_c[idx];
    double a00 = unwrap_angle(ac[i00], ca);
    double a01 = unwrap_angle(ac[i01], ca);
    double a10 = unwrap_angle(ac[i10], ca);
    double a11 = unwrap_angle(ac[i11], ca);
    double poly[4][2] = {{r00, a00}, {r01, a01}, {r11, a11}, {r10, a10}};
    double area_total = fabs(polygon_area((const double(*)[2])poly, 4));
    if (area_total <= 0.0) {
      double pos = (radial_c[idx] - min) / step;
      if (pos < 0.0) continue;
      long long bin = sat_isize(floor(pos));
      if (bin >= 0 && (unsigned long long)bin < n) bucket_push(&bk[bin], idx, 1.0);
      continue;
    }
    double rmin = r00, rmax = r00;
    double rv[3] = {r01, r10, r11};
    for (int k = 0; k < 3; k++) {
      if (rv[k] < rmin) rmin = rv[k];
      if (rv[k] > rmax) rmax = rv[k];
    }
    long long start = sat_isize(floor((rmin - min) / step));
    long long end = sat_isize(floor((rmax - min) / step));
    if (end < 0 || start >= (long long)n) continue;
    if (start < 0) start = 0;
    if (end >= (long long)n) end = (long long)n - 1;
    double amin = a00, amax = a00;
    double av[3] = {a01, a10, a11};
    for (int k = 0; k < 3; k++) {
      if (av[k] < amin) amin = av[k];
      if (av[k] > amax) amax = av[k];
    }
    for (long long bb = start; bb <= end; bb++) {
      double bin_min = min + (double)bb * step;
      double bin_max = bin_min + step;
      double clipped[24][2];
      int m = clip_rect((const double(*)[2])poly, 4, bin_min, bin_max, amin, amax, clipped);
      double area = fabs(polygon_area((const double(*)[2])clipped, m));
      if (area > 0.0) bucket_push(&bk[bb], idx, area / area_total);
    }
  }
}

/* pixel_split.rs:389-455 (build_lut1d_ps_csr and _with_range). */
int or_build_lut1d_ps(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, int unit,
                      int has_range, double rmin, double rmax, or_lut *out) {
  if (n_points == 0 || dim1 == 0 || dim2 == 0) return 1;
  size_t nc = (dim1 + 1) * (dim2 + 1), npix = dim1 * dim2;
  double *tc = (double *)malloc(nc * sizeof(double));
  double *ac = (double *)malloc(nc * sizeof(double));
  double *rc = (double *)malloc(nc * sizeof(double));
  double *tth = (double *)malloc(npix * sizeof(double));
  double *chi = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  or_corners(p, dim1, dim2, tc, ac);
  or_convert_unit(unit, p->wavelength, tc, nc, rc);
  or_centres(p, dim1, dim2, tth, chi);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  double min, max, step;
  if (has_range) {
    /* histogram.rs:184-188 */
    min = rmin;
    max = rmax + 10.0 * OR_EPS;
  } else {
    min = vmin(rc, nc);
    max = vmax(rc, nc);
    max += 10.0 * OR_EPS;
  }
  step = (max - min) / (double)n_points;
  bucket *bk = (bucket *)calloc(n_points, sizeof(bucket));
  ps_fractions(p, dim1, dim2, rad, chi, unit, n_points, min, step, rc, ac, !has_range, bk);
  buckets_to_lut(bk, n_points, 1, min, step, out);
  free(bk);
  free(tc);
  free(ac);
  free(rc);
  free(tth);
  free(chi);
  free(rad);
  return 0;
}

/* counting_csr.rs:54-134 (build_lut1d_csr and _with_range) on given radial
 * values, with the bin rules of histogram.rs:126-141, 184-196. */
int or_counting_from_radial(const double *rad, size_t npix, size_t n_points, int has_range,
                            double rmin, double rmax, or_lut *out) {
  if (n_points == 0 || npix == 0) return 1;
  double min, max, step;
  if (has_range) {
    min = rmin;
    max = rmax + 10.0 * OR_EPS;
  } else {
    min = vmin(rad, npix);
    max = vmax(rad, npix) + 10.0 * OR_EPS;
  }
  step = (max - min) / (double)n_points;
  bucket *bk = (bucket *)calloc(n_points, sizeof(bucket));
  for (size_t i = 0; i < npix; i++) {
    double r = rad[i];
    if (has_range && !(r >= min && r <= max)) continue;
    unsigned long long bin = sat_usize((r - min) / step);
    if (bin > n_points - 1) bin = n_points - 1;
    bucket_push(&bk[bin], i, 1.0);
  }
  buckets_to_lut(bk, n_points, 0, min, step, out);
  free(bk);
  return 0;
}

int or_build_lut1d_counting(const or_poni *p, size_t dim1, size_t dim2, size_t n_points,
                            int unit, int has_range, double rmin, double rmax, or_lut *out) {
  if (n_points == 0 || dim1 == 0 || dim2 == 0) return 1;
  size_t npix = dim1 * dim2;
  double *tth = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  or_centres(p, dim1, dim2, tth, NULL);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  int rc = or_counting_from_radial(rad, npix, n_points, has_range, rmin, rmax, out);
  free(tth);
  free(rad);
  return rc;
}

/* pixel_split.rs:190-328 on an explicit (min, step), as the reference's
 * tests call it (lut1d_ps_fractions). */
int or_ps_fractions(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, int unit,
                    double min, double step, int clamp_low, or_lut *out) {
  if (n_points == 0 || dim1 == 0 || dim2 == 0) return 1;
  size_t nc = (dim1 + 1) * (dim2 + 1), npix = dim1 * dim2;
  double *tc = (double *)malloc(nc * sizeof(double));
  double *ac = (double *)malloc(nc * sizeof(double));
  double *rc = (double *)malloc(nc * sizeof(double));
  double *tth = (double *)malloc(npix * sizeof(double));
  double *chi = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  or_corners(p, dim1, dim2, tc, ac);
  or_convert_unit(unit, p->wavelength, tc, nc, rc);
  or_centres(p, dim1, dim2, tth, chi);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  bucket *bk = (bucket *)calloc(n_points, sizeof(bucket));
  ps_fractions(p, dim1, dim2, rad, chi, unit, n_points, min, step, rc, ac, clamp_low, bk);
  buckets_to_lut(bk, n_points, 1, min, step, out);
  free(bk);
  free(tc);
  free(ac);
  free(rc);
  free(tth);
  free(chi);
  free(rad);
  return 0;
}

/* ------------------------------------------------------------- 2D grids */

/* integration/grid.rs:26-222 (IntegrationGrid2d). */
typedef struct {
  size_t nr, na;
  double r0, r1, a0, a1;
  int clamp_radial_end, clamp_index_rounding, clamp_radial_outside, clamp_azimuthal_outside;
} grid2d;

/* grid.rs:48-108: validity checks and the correction of distant starts. */
static int grid_new(grid2d *g, size_t nr, size_t na, double rlo, double rhi, double alo,
                    double ahi) {
  memset(g, 0, sizeof(*g));
  if (nr == 0 || na == 0) return 1;
  double lo[2] = {rlo, alo}, hi[2] = {rhi, ahi};
  size_t cnt[2] = {nr, na};
  for (int k = 0; k < 2; k++) {
    double step = (hi[k] - lo[k]) / (double)cnt[k];
    if (!isfinite(lo[k]) || !isfinite(hi[k]) || !isfinite(step) || step <= 0.0 ||
        lo[k] + step == lo[k] || hi[k] - step == hi[k])
      return 1;
  }
  if (rlo < 0.0) return 1;
  if (ahi - alo > OR_TAU + 1e-12) return 1;
  if (!(alo >= -OR_TAU && alo <= OR_TAU)) {
    double start = rem_euclid(alo, OR_TAU);
    if (start >= OR_PI) start -= OR_TAU;
    ahi = start + (ahi - alo);
    alo = start;
  }
  g->nr = nr;
  g->na = na;
  g->r0 = rlo;
  g->r1 = rhi;
  g->a0 = alo;
  g->a1 = ahi;
  return 0;
}

static double grid_rstep(const grid2d *g) { return (g->r1 - g->r0) / (double)g->nr; }
static double grid_astep(const grid2d *g) { return (g->a1 - g->a0) / (double)g->na; }

/* grid.rs:176-221 (IntegrationGrid2d::bin); -1 for no bin. */
static long long grid_bin(const grid2d *g, double radial, double az) {
  double r0 = g->r0, r1 = g->r1, a0 = g->a0, a1 = g->a1;
  if (!isfinite(radial) || !isfinite(az)) return -1;
  if (!g->clamp_radial_outside &&
      (radial < r0 || radial > r1 || (radial == r1 && !g->clamp_radial_end)))
    return -1;
  double offset;
  if (g->clamp_azimuthal_outside) {
    offset = az - a0;
    double hi = a1 - a0;
    offset = offset < 0.0 ? 0.0 : (offset > hi ? hi : offset);
  } else if (g->clamp_index_rounding) {
    double wrapped = (az >= a0 && az < a1) ? az : a0 + rem_euclid(az - a0, OR_TAU);
    if (wrapped >= a1 && a1 - a0 < OR_TAU) return -1;
    offset = wrapped - a0;
  } else {
    double a = a0 + rem_euclid(az - a0, OR_TAU);
    if (a >= a1) return -1;
    offset = a - a0;
  }
  unsigned long long r = sat_usize((radial - r0) / grid_rstep(g));
  unsigned long long a = sat_usize(offset / grid_astep(g));
  if (g->clamp_radial_end || g->clamp_index_rounding || g->clamp_radial_outside)
    if (r > g->nr - 1) r = g->nr - 1;
  if (g->clamp_index_rounding || g->clamp_azimuthal_outside)
    if (a > g->na - 1) a = g->na - 1;
  if (r < g->nr && a < g->na) return (long long)(a * g->nr + r);
  return -1;
}

/* shared_grid.rs:202-232: 128 x 128 x 4 samples of a beam pixel on a 2D grid;
 * weight 1/65536 per sample, summed per bin (sums of 2^-16 are exact). */
static void sample_pixel_2d(const or_poni *p, const trig *t, size_t dim2, size_t pixel, int unit,
                            const grid2d *g, bucket *bk) {
  const int N = 128;
  const double off[4][2] = {{0.25, 0.5}, {0.5, 0.25}, {0.75, 0.5}, {0.5, 0.75}};
  size_t nb = g->nr * g->na;
  unsigned *hits = (unsigned *)calloc(nb, sizeof(unsigned));
  size_t row = pixel / dim2, col = pixel % dim2;
  double cv[1];
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int o = 0; o < 4; o++) {
        double d1 = (double)row + ((double)i + off[o][0]) / (double)N;
        double d2 = ((double)(col * N + j) + off[o][1]) / (double)N;
        double x, y, z;
        point_xyz(p, t, d1, d2, &x, &y, &z);
        double tth = f_tth(x, y, z);
        double az = f_azimuth(x, y);
        or_convert_unit(unit, p->wavelength, &tth, 1, cv);
        long long b = grid_bin(g, cv[0], az);
        if (b >= 0) hits[b]++;
      }
  for (size_t b = 0; b < nb; b++)
    if (hits[b]) bucket_push(&bk[b], pixel, (double)hits[b] / 65536.0);
  free(hits);
}

/* shared_grid.rs:152-200 (split_polygon). Duplicate bins of the periodic
 * copies are merged by adding in emission order (shift, ar, rr). */
static void split_polygon_2d(const double (*poly)[2], double area, const grid2d *g, size_t pixel,
                             bucket *bk) {
  double r0 = g->r0, a0 = g->a0, dr = grid_rstep(g), da = grid_astep(g);
  double rmin = INFINITY, rmax = -INFINITY;
  for (int i = 0; i < 4; i++) {
    rmin = fmin(rmin, poly[i][0]);
    rmax = fmax(rmax, poly[i][0]);
  }
  long long rs = sat_isize(floor((rmin - r0) / dr));
  long long re = sat_isize(floor((rmax - r0) / dr));
  if (rs < 0) rs = 0;
  if (re > (long long)g->nr - 1) re = (long long)g->nr - 1;
  size_t cap = 16, len = 0;
  size_t *bins = (size_t *)malloc(cap * sizeof(size_t));
  double *ws = (double *)malloc(cap * sizeof(double));
  const double shifts[3] = {-OR_TAU, 0.0, OR_TAU};
  for (int s = 0; s < 3; s++) {
    double ps[4][2];
    double amin = INFINITY, amax = -INFINITY;
    for (int i = 0; i < 4; i++) {
      ps[i][0] = poly[i][0];
      ps[i][1] = poly[i][1] + shifts[s];
      amin = fmin(amin, ps[i][1]);
      amax = fmax(amax, ps[i][1]);
    }
    long long as = sat_isize(floor((amin - a0) / da));
    long long ae = sat_isize(floor((amax - a0) / da));
    if (as < 0) as = 0;
    if (ae > (long long)g->na - 1) ae = (long long)g->na - 1;
    for (long long ar = as; ar <= ae; ar++)
      for (long long rr = rs; rr <= re; rr++) {
        double r = r0 + (double)rr * dr;
        double a = a0 + (double)ar * da;
        double clipped[24][2];
        int m = clip_rect((const double(*)[2])ps, 4, r, r + dr, a, a + da, clipped);
        double w = fabs(polygon_area((const double(*)[2])clipped, m)) / area;
        if (w > 0.0) {
          size_t bin = (size_t)ar * g->nr + (size_t)rr;
          size_t k;
          for (k = 0; k < len; k++)
            if (bins[k] == bin) break;
          if (k < len) {
            ws[k] += w;
          } else {
            if (len == cap) {
              cap *= 2;
              bins = (size_t *)realloc(bins, cap * sizeof(size_t));
              ws = (double *)realloc(ws, cap * sizeof(double));
            }
            bins[len] = bin;
            ws[len] = w;
            len++;
          }
        }
      }
  }
  for (size_t k = 0; k < len; k++) bucket_push(&bk[bins[k]], pixel, ws[k]);
  free(bins);
  free(ws);
}

/* integration/methods.rs:365-388, grid.rs:579-675, 978-1028 and
 * shared_grid.rs:13-150: the default 2D table of Counting and PixelSplit
 * (or an explicit radial and/or azimuthal range). */
int or_build_lut2d(const or_poni *p, size_t dim1, size_t dim2, int method, int unit,
                   size_t n_radial, size_t n_azimuthal, int has_rr, double rlo, double rhi,
                   int has_ar, double alo, double ahi, or_lut *out, double *edges) {
  if (dim1 == 0 || dim2 == 0 || (method != 0 && method != 1)) return 1;
  size_t nc = (dim1 + 1) * (dim2 + 1), npix = dim1 * dim2;
  double *tth = (double *)malloc(npix * sizeof(double));
  double *chi = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  or_centres(p, dim1, dim2, tth, chi);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  double *tc = NULL, *ac = NULL, *rc = NULL;
  if (method == 1) {
    tc = (double *)malloc(nc * sizeof(double));
    ac = (double *)malloc(nc * sizeof(double));
    rc = (double *)malloc(nc * sizeof(double));
    or_corners(p, dim1, dim2, tc, ac);
    or_convert_unit(unit, p->wavelength, tc, nc, rc);
  }
  /* default radial edges (grid.rs:978-1028) */
  double r_lo = rlo, r_hi = rhi;
  if (!has_rr) {
    if (method == 0) {
      r_lo = vmin(rad, npix);
      r_hi = vmax(rad, npix) + 10.0 * OR_EPS;
    } else {
      r_lo = vmin(rc, nc);
      r_hi = vmax(rc, nc) + 10.0 * OR_EPS;
    }
  }
  /* default azimuthal edges (grid.rs:629-675) */
  double a_lo = alo, a_hi = ahi;
  if (!has_ar) {
    if (method == 0) {
      a_lo = vmin(chi, npix);
      a_hi = vmax(chi, npix) + 10.0 * OR_EPS;
    } else {
      a_lo = vmin(ac, nc);
      a_hi = vmax(ac, nc) + 10.0 * OR_EPS;
    }
    if (!(a_hi - a_lo > 1e-12)) {
      a_lo = -OR_PI;
      a_hi = OR_PI;
    }
  }
  grid2d g;
  if (grid_new(&g, n_radial, n_azimuthal, r_lo, r_hi, a_lo, a_hi)) {
    free(tth); free(chi); free(rad); free(tc); free(ac); free(rc);
    return 2;
  }
  g.clamp_index_rounding = 1;
  if (!has_rr && method == 0) g.clamp_radial_end = 1;
  if (edges) {
    edges[0] = g.r0; edges[1] = g.r1; edges[2] = g.a0; edges[3] = g.a1;
  }
  trig t = make_trig(p);
  size_t nb = n_radial * n_azimuthal;
  bucket *bk = (bucket *)calloc(nb, sizeof(bucket));
  size_t w = dim2 + 1;
  for (size_t pix = 0; pix < npix; pix++) {
    if (method == 0) {
      long long b = grid_bin(&g, rad[pix], chi[pix]);
      if (b >= 0) bucket_push(&bk[b], pix, 1.0);
      continue;
    }
    size_t b0 = pix / dim2 * w + pix % dim2;
    size_t idx[4] = {b0, b0 + 1, b0 + w + 1, b0 + w};
    int axis_on_edge = 0, zero_r = 0;
    double winding = 0.0;
    for (int i = 0; i < 4; i++) {
      double delta = ac[idx[(i + 1) % 4]] - ac[idx[i]];
      delta = rem_euclid(delta + OR_PI, OR_TAU) - OR_PI;
      if (fabs(delta) >= OR_PI - 64.0 * OR_EPS) axis_on_edge = 1;
      winding += delta;
      if (rc[idx[i]] == 0.0) zero_r = 1;
    }
    if (fabs(winding) > OR_PI || axis_on_edge || zero_r) {
      sample_pixel_2d(p, &t, dim2, pix, unit, &g, bk);
      continue;
    }
    double center = g.a0 + rem_euclid(chi[pix] - g.a0, OR_TAU);
    double poly[4][2];
    for (int i = 0; i < 4; i++) {
      poly[i][0] = rc[idx[i]];
      poly[i][1] = center + rem_euclid(ac[idx[i]] - center + OR_PI, OR_TAU) - OR_PI;
    }
    double area = fabs(polygon_area((const double(*)[2])poly, 4));
    if (area > 0.0) {
      split_polygon_2d((const double(*)[2])poly, area, &g, pix, bk);
    } else {
      long long b = grid_bin(&g, rad[pix], chi[pix]);
      if (b >= 0) bucket_push(&bk[b], pix, 1.0);
    }
  }
  buckets_to_lut(bk, nb, method == 1, g.r0, grid_rstep(&g), out);
  free(bk);
  free(tth); free(chi); free(rad); free(tc); free(ac); free(rc);
  return 0;
}

/* ----------------------------------------------------------- supersampling */

/* lut/common.rs:33-65 (hull_grid): the outermost subpixel centres. */
static size_t hull_grid(size_t dim1, size_t dim2, size_t n_ss, double **d1o, double **d2o) {
  size_t cap = 2 * dim1 * n_ss + 2 * dim2 * n_ss;
  double *d1 = (double *)malloc(cap * sizeof(double));
  double *d2 = (double *)malloc(cap * sizeof(double));
  double border = 0.5 / (double)n_ss, step = 1.0 / (double)n_ss;
  size_t k = 0;
  for (size_t i = 0; i < dim1 * n_ss; i++) { d1[k] = border + (double)i * step; d2[k++] = border; }
  for (size_t i = 0; i < dim1 * n_ss; i++) {
    d1[k] = border + (double)i * step;
    d2[k++] = (double)dim2 - border;
  }
  for (size_t i = 1; i < dim2 * n_ss; i++) { d1[k] = border; d2[k++] = border + (double)i * step; }
  for (size_t i = 1; i < dim2 * n_ss; i++) {
    d1[k] = (double)dim1 - border;
    d2[k++] = border + (double)i * step;
  }
  *d1o = d1;
  *d2o = d2;
  return k;
}

/* histogram.rs:151-181: range over pixel centres and the hull grid, + 3 eps. */
static void ss_ranges(const or_poni *p, size_t dim1, size_t dim2, size_t n_ss, int unit,
                      const double *rad, const double *chi, size_t npix, double *rlo,
                      double *rhi, double *alo, double *ahi) {
  double *hd1, *hd2;
  size_t nh = hull_grid(dim1, dim2, n_ss, &hd1, &hd2);
  double *ht = (double *)malloc((nh ? nh : 1) * sizeof(double));
  double *ha = (double *)malloc((nh ? nh : 1) * sizeof(double));
  double *hr = (double *)malloc((nh ? nh : 1) * sizeof(double));
  or_tth_chi(p, hd1, hd2, nh, ht, ha);
  or_convert_unit(unit, p->wavelength, ht, nh, hr);
  *rlo = fmin(vmin(rad, npix), vmin(hr, nh));
  *rhi = fmax(vmax(rad, npix), vmax(hr, nh)) + 3.0 * OR_EPS;
  if (chi) {
    *alo = fmin(vmin(chi, npix), vmin(ha, nh));
    *ahi = fmax(vmax(chi, npix), vmax(ha, nh)) + 3.0 * OR_EPS;
  }
  free(hd1); free(hd2); free(ht); free(ha); free(hr);
}

/* Per-pixel subpixel counts (supersampled.rs:21-106, 171-500, 731-1379):
 * every subpixel is evaluated exactly; `inv` selects the fast builders'
 * coordinates rf + (i + 0.5) * inv_ss (else (i + 0.5) / n_ss). A weight is
 * count / n_ss^2. */
static void ss_count(const or_poni *p, size_t dim1, size_t dim2, size_t n_ss, int unit, int inv,
                     int twod, size_t n1, double min1, double max1, double step1, int clamped1,
                     const grid2d *g, bucket *bk) {
  trig t = make_trig(p);
  double inv_ss = 1.0 / (double)n_ss;
  size_t cap = n_ss * n_ss;
  size_t *bins = (size_t *)malloc(cap * sizeof(size_t));
  size_t *cnts = (size_t *)malloc(cap * sizeof(size_t));
  double cv[1];
  for (size_t row = 0; row < dim1; row++)
    for (size_t col = 0; col < dim2; col++) {
      size_t len = 0;
      for (size_t i = 0; i < n_ss; i++) {
        double d1 = inv ? (double)row + ((double)i + 0.5) * inv_ss
                        : (double)row + ((double)i + 0.5) / (double)n_ss;
        for (size_t j = 0; j < n_ss; j++) {
          double d2 = inv ? ((double)(col * n_ss + j) + 0.5) * inv_ss
                          : ((double)(col * n_ss + j) + 0.5) / (double)n_ss;
          double x, y, z;
          point_xyz(p, &t, d1, d2, &x, &y, &z);
          double tth = f_tth(x, y, z);
          or_convert_unit(unit, p->wavelength, &tth, 1, cv);
          double r = cv[0];
          long long b;
          if (twod) {
            b = grid_bin(g, r, f_azimuth(x, y));
          } else if (clamped1) {
            unsigned long long u = sat_usize((r - min1) / step1);
            b = (long long)(u > n1 - 1 ? n1 - 1 : u);
          } else {
            if (!(r >= min1 && r <= max1)) continue;
            unsigned long long u = sat_usize((r - min1) / step1);
            b = (long long)(u > n1 - 1 ? n1 - 1 : u);
          }
          if (b < 0) continue;
          size_t k;
          for (k = 0; k < len; k++)
            if (bins[k] == (size_t)b) break;
          if (k < len) cnts[k]++;
          else { bins[len] = (size_t)b; cnts[len] = 1; len++; }
        }
      }
      for (size_t k = 0; k < len; k++)
        bucket_push(&bk[bins[k]], row * dim2 + col, (double)cnts[k] / (double)(n_ss * n_ss));
    }
  free(bins);
  free(cnts);
}

/* supersampled.rs:517-595 (build_lut1d_ss_csr and _with_range). */
int or_build_lut1d_ss(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, size_t n_ss,
                      int unit, int has_range, double rmin, double rmax, or_lut *out) {
  if (n_points == 0 || dim1 == 0 || dim2 == 0 || n_ss == 0) return 1;
  size_t npix = dim1 * dim2;
  double *tth = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  or_centres(p, dim1, dim2, tth, NULL);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  double min, max, step, dummy;
  if (has_range) {
    min = rmin;
    max = rmax + 10.0 * OR_EPS;
  } else {
    ss_ranges(p, dim1, dim2, n_ss, unit, rad, NULL, npix, &min, &max, &dummy, &dummy);
  }
  step = (max - min) / (double)n_points;
  bucket *bk = (bucket *)calloc(n_points, sizeof(bucket));
  ss_count(p, dim1, dim2, n_ss, unit, 1, 0, n_points, min, max, step, !has_range, NULL, bk);
  buckets_to_lut(bk, n_points, 1, min, step, out);
  free(bk);
  free(tth);
  free(rad);
  return 0;
}

/* Default (or explicit-range) supersampled 2D table: methods.rs:365-388,
 * grid.rs:579-675, 978-1028, supersampled.rs:1437-1477. */
int or_build_lut2d_ss(const or_poni *p, size_t dim1, size_t dim2, size_t n_radial,
                      size_t n_azimuthal, size_t n_ss, int unit, int has_rr, double rlo,
                      double rhi, int has_ar, double alo, double ahi, or_lut *out,
                      double *edges) {
  if (dim1 == 0 || dim2 == 0 || n_ss == 0) return 1;
  size_t npix = dim1 * dim2;
  double *tth = (double *)malloc(npix * sizeof(double));
  double *chi = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  or_centres(p, dim1, dim2, tth, chi);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  double s_rlo, s_rhi, s_alo, s_ahi;
  ss_ranges(p, dim1, dim2, n_ss, unit, rad, chi, npix, &s_rlo, &s_rhi, &s_alo, &s_ahi);
  double r_lo = has_rr ? rlo : s_rlo, r_hi = has_rr ? rhi : s_rhi;
  double a_lo = alo, a_hi = ahi;
  if (!has_ar) {
    a_lo = s_alo;
    a_hi = s_ahi;
    if (!(a_hi - a_lo > 1e-12)) {
      a_lo = -OR_PI;
      a_hi = OR_PI;
    }
  }
  grid2d g;
  if (grid_new(&g, n_radial, n_azimuthal, r_lo, r_hi, a_lo, a_hi)) {
    free(tth); free(chi); free(rad);
    return 2;
  }
  g.clamp_index_rounding = 1;
  if (!has_rr) g.clamp_radial_outside = 1;
  if (!has_ar) g.clamp_azimuthal_outside = 1;
  int fast = g.clamp_radial_outside && g.clamp_azimuthal_outside && grid_astep(&g) > 1e-9;
  if (edges) { edges[0] = g.r0; edges[1] = g.r1; edges[2] = g.a0; edges[3] = g.a1; }
  size_t nb = n_radial * n_azimuthal;
  bucket *bk = (bucket *)calloc(nb, sizeof(bucket));
  ss_count(p, dim1, dim2, n_ss, unit, fast, 1, 0, 0.0, 0.0, 0.0, 0, &g, bk);
  buckets_to_lut(bk, nb, 1, g.r0, grid_rstep(&g), out);
  free(bk);
  free(tth); free(chi); free(rad);
  return 0;
}

/* -------------------------------------------------- pyFAI-compatible split */

/* pyfai_1d.rs:36-46: [max(min corner, 0), max corner * (1 + f32::EPSILON)]. */
static void pyfai_range(const double *rc, size_t nc, double *lo, double *hi) {
  double mn = INFINITY, mx = -INFINITY;
  for (size_t i = 0; i < nc; i++) {
    mn = fmin(mn, rc[i]);
    mx = fmax(mx, rc[i]);
  }
  *lo = fmax(mn, 0.0);
  *hi = mx * (1.0 + (double)FLT_EPSILON);
}

/* pyfai_1d.rs:259-354 (split1d): the azimuthal range [a0, a1] of the grid
 * sets the seam offset; with `sector`, a pixel whose corners all lie below
 * a0 or all above a1 is dropped; `sign` is corner_sign (:131-154; +1 with no
 * detector). Appends the pixel's entries. */
static void split1d(double poly[4][2], double r0, double r1, size_t bins, double a0, double a1,
                    int sector, double sign, size_t pixel, bucket *bk) {
  double offset = (a0 >= -OR_PI && a1 <= OR_PI) ? 0.0
                  : (a0 >= 0.0 && a1 <= OR_TAU) ? OR_PI : a0 + OR_PI;
  for (int i = 0; i < 4; i++) poly[i][1] = wrap_tau(poly[i][1] - offset + OR_PI) - OR_PI;
  double *a = poly[0], *b = poly[1], *c = poly[2], *d = poly[3];
  double area = sign * (0.5 * (c[0] - a[0]) * (d[1] - b[1]) - (c[1] - a[1]) * (d[0] - b[0]));
  if (area > 0.0) {
    for (int i = 0; i < 4; i++)
      if (poly[i][1] < 0.0) poly[i][1] += OR_TAU;
    double s = 0.0;
    s = s + poly[0][1];
    s = s + poly[1][1];
    s = s + poly[2][1];
    s = s + poly[3][1];
    if (s * 0.25 > OR_PI)
      for (int i = 0; i < 4; i++) poly[i][1] -= OR_TAU;
  }
  for (int i = 0; i < 4; i++) poly[i][1] += offset;
  if (sector) {
    int below = 1, above = 1;
    for (int i = 0; i < 4; i++) {
      below = below && poly[i][1] < a0;
      above = above && poly[i][1] > a1;
    }
    if (below || above) return;
  }
  double dr = (r1 - r0) / (double)bins;
  for (int i = 0; i < 4; i++) poly[i][0] = (poly[i][0] - r0) / dr;
  double lo = INFINITY, hi = -INFINITY;
  for (int i = 0; i < 4; i++) {
    lo = fmin(lo, poly[i][0]);
    hi = fmax(hi, poly[i][0]);
  }
  if (hi < 0.0 || lo >= (double)bins) return;
  size_t first = (size_t)sat_usize(fmax(floor(lo), 0.0));
  long long e = sat_isize(floor(hi)) + 1;
  if (e < 0) e = 0;
  if (e > (long long)bins) e = (long long)bins;
  size_t end = (size_t)e;
  if (floor(lo) == floor(hi)) {
    bucket_push(&bk[first], pixel, 1.0);
    return;
  }
  size_t n = end > first ? end - first : 0;
  float *buf = (float *)calloc(n ? n : 1, sizeof(float));
  for (int i = 0; i < 4; i++) {
    double x0 = poly[i][0], y0 = poly[i][1], x1 = poly[(i + 1) % 4][0], y1 = poly[(i + 1) % 4][1];
    if (x0 == x1) continue;
    double slope = (y1 - y0) / (x1 - x0);
    double intercept = y0 - slope * x0;
    double sign = x1 > x0 ? 1.0 : -1.0;
    for (size_t j = 0; j < n; j++) {
      double left = fmax(fmin(x0, x1), (double)(first + j));
      double right = fmin(fmax(x0, x1), (double)(first + j + 1));
      if (right > left)
        buf[j] = buf[j] + (float)(sign * (right - left) * (0.5 * slope * (right + left) + intercept));
    }
  }
  double sum = 0.0;
  for (size_t j = 0; j < n; j++) sum += (double)buf[j];
  if (!isfinite(sum) || sum == 0.0) {
    free(buf);
    return;
  }
  double total = 0.0;
  for (size_t j = 0; j < n; j++) {
    double w = (double)buf[j] / sum;
    if (isfinite(w) && w > 0.0) total += w;
  }
  for (size_t j = 0; j < n; j++) {
    double w = (double)buf[j] / sum;
    if (isfinite(w) && w > 0.0) bucket_push(&bk[first + j], pixel, w / total);
  }
  free(buf);
}

/* pyfai_1d.rs:55-109 (build_lut1d_pyfai_csr) with the full azimuth. */
int or_build_lut1d_pyfai(const or_poni *p, size_t dim1, size_t dim2, size_t n_points, int unit,
                         int has_range, double rmin, double rmax, or_lut *out) {
  if (n_points == 0 || dim1 == 0 || dim2 == 0) return 1;
  size_t nc = (dim1 + 1) * (dim2 + 1), npix = dim1 * dim2;
  double *tc = (double *)malloc(nc * sizeof(double));
  double *ac = (double *)malloc(nc * sizeof(double));
  double *rc = (double *)malloc(nc * sizeof(double));
  or_corners(p, dim1, dim2, tc, ac);
  or_convert_unit(unit, p->wavelength, tc, nc, rc);
  double r0 = rmin, r1 = rmax;
  if (!has_range) pyfai_range(rc, nc, &r0, &r1);
  if (!sample_grid_ok(n_points, r0, r1)) {
    free(tc); free(ac); free(rc);
    return 2;
  }
  trig t = make_trig(p);
  bucket *bk = (bucket *)calloc(n_points, sizeof(bucket));
  size_t w = dim2 + 1, entries = 0;
  for (size_t pix = 0; pix < npix; pix++) {
    size_t b = pix / dim2 * w + pix % dim2;
    size_t idx[4] = {b, b + w, b + w + 1, b + 1}; /* pyFAI order a, d, c, b */
    double poly[4][2];
    for (int i = 0; i < 4; i++) {
      poly[i][0] = rc[idx[i]];
      poly[i][1] = ac[idx[i]];
    }
    if (singular((const double(*)[2])poly)) sample_pixel(p, &t, dim2, pix, unit, n_points, r0, r1,
                                                          !has_range, bk);
    else split1d(poly, r0, r1, n_points, -OR_PI, OR_PI, 0, 1.0, pix, bk);
  }
  for (size_t k = 0; k < n_points; k++) entries += bk[k].len;
  buckets_to_lut(bk, n_points, 1, r0, (r1 - r0) / (double)n_points, out);
  free(bk);
  free(tc); free(ac); free(rc);
  return entries > 64u * 1024u * 1024u ? 3 : 0;
}

/* pyfai_split.rs:139-167 (line): signed strip areas distributed down the
 * azimuthal rows of each radial column, in float32. */
static void pyfai_line(float *buf, size_t stride, const double *start, const double *stop) {
  if (start[0] == stop[0]) return;
  float slope = (float)((stop[1] - start[1]) / (stop[0] - start[0]));
  float intercept = (float)(stop[1] - (double)slope * stop[0]);
  double lo = fmin(start[0], stop[0]), hi = fmax(start[0], stop[0]);
  double sign = stop[0] > start[0] ? 1.0 : -1.0;
  size_t r_end = (size_t)sat_usize(ceil(hi));
  for (size_t r = (size_t)sat_usize(floor(lo)); r < r_end; r++) {
    double left = fmax(lo, (double)r), right = fmin(hi, (double)(r + 1));
    float area = (float)(sign * (right - left) *
                         (0.5 * (double)slope * (right + left) + (double)intercept));
    float width = (float)(right - left);
    float remaining = fabsf(area);
    for (size_t a = 0; a < stride; a++) {
      if (remaining <= 0.0f) break;
      float inc = fminf(width, remaining);
      buf[r * stride + a] = buf[r * stride + a] + copysignf(inc, area);
      remaining = remaining - inc;
    }
  }
}

static double clampd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }

/* pyfai_split.rs:35-137 (split) with sign = +1. Returns 0 when the caller
 * must use the fractional fallback (None), else 1 after appending entries. */
static int pyfai_split2d(double poly[4][2], const grid2d *g, size_t pixel, bucket *bk) {
  double a0 = g->a0, a1 = g->a1;
  double cstart = (a0 >= -OR_PI && a1 <= OR_PI) ? -OR_PI : (a0 >= 0.0 && a1 <= OR_TAU) ? 0.0 : a0;
  double offset = cstart + OR_PI;
  for (int i = 0; i < 4; i++) poly[i][1] = wrap_tau(poly[i][1] - offset + OR_PI) - OR_PI;
  double *a = poly[0], *b = poly[1], *c = poly[2], *d = poly[3];
  double area = 1.0 * (0.5 * (c[0] - a[0]) * (d[1] - b[1]) - (c[1] - a[1]) * (d[0] - b[0]));
  if (area > 0.0) {
    for (int i = 0; i < 4; i++)
      if (poly[i][1] < 0.0) poly[i][1] += OR_TAU;
    double s = 0.0;
    for (int i = 0; i < 4; i++) s = s + poly[i][1];
    if (s * 0.25 > OR_PI)
      for (int i = 0; i < 4; i++) poly[i][1] -= OR_TAU;
  }
  for (int i = 0; i < 4; i++) poly[i][1] += offset;
  double r0 = g->r0, r1 = g->r1;
  int all_lo = 1, all_hi = 1, all_alo = 1, all_ahi = 1;
  for (int i = 0; i < 4; i++) {
    all_lo &= poly[i][0] < r0;
    all_hi &= poly[i][0] > r1;
    all_alo &= poly[i][1] < a0;
    all_ahi &= poly[i][1] > a1;
  }
  if (all_lo || all_hi || all_alo || all_ahi) return 1;
  double dr = grid_rstep(g), da = grid_astep(g);
  for (int i = 0; i < 4; i++) {
    poly[i][0] = (clampd(poly[i][0], r0, r1) - r0) / dr;
    poly[i][1] = (clampd(poly[i][1], a0, a1) - a0) / da;
  }
  double mr = INFINITY, ma = INFINITY;
  for (int i = 0; i < 4; i++) {
    mr = fmin(mr, poly[i][0]);
    ma = fmin(ma, poly[i][1]);
  }
  size_t ir = (size_t)sat_usize(floor(mr)), ia = (size_t)sat_usize(floor(ma));
  for (int i = 0; i < 4; i++) {
    poly[i][0] -= (double)ir;
    poly[i][1] -= (double)ia;
  }
  double xr = 0.0, xa = 0.0;
  for (int i = 0; i < 4; i++) {
    xr = fmax(xr, poly[i][0]);
    xa = fmax(xa, poly[i][1]);
  }
  size_t nr = (size_t)sat_usize(ceil(xr)), na = (size_t)sat_usize(ceil(xa));
  if (nr == 0 || na == 0) return 0;
  float *buf = (float *)calloc((nr + 1) * (na + 1), sizeof(float));
  for (int i = 0; i < 4; i++) pyfai_line(buf, na + 1, poly[i], poly[(i + 1) % 4]);
  double total = 0.0;
  size_t cnt = 0;
  for (size_t r = 0; r < nr; r++)
    for (size_t q = 0; q < na; q++) {
      double v = 1.0 * (double)buf[r * (na + 1) + q];
      if (!isfinite(v)) {
        free(buf);
        return 0;
      }
      if (v > 0.0 && ir + r < g->nr && ia + q < g->na) {
        total += v;
        cnt++;
      }
    }
  if (total <= 0.0) {
    free(buf);
    return 0;
  }
  for (size_t r = 0; r < nr; r++)
    for (size_t q = 0; q < na; q++) {
      double v = 1.0 * (double)buf[r * (na + 1) + q];
      if (v > 0.0 && ir + r < g->nr && ia + q < g->na)
        bucket_push(&bk[(ia + q) * g->nr + ir + r], pixel, v / total);
    }
  free(buf);
  return 1;
}

/* pyFAI-compatible 2D table (shared_grid.rs:13-150 with
 * PixelSplitMode::PyfaiHistogram): grid chi (-pi, pi) by default
 * (methods.rs:371-373), radial edges from pyfai_radial_range. */
int or_build_lut2d_pyfai(const or_poni *p, size_t dim1, size_t dim2, int unit, size_t n_radial,
                         size_t n_azimuthal, int has_rr, double rlo, double rhi, int has_ar,
                         double alo, double ahi, or_lut *out, double *edges) {
  if (dim1 == 0 || dim2 == 0) return 1;
  size_t nc = (dim1 + 1) * (dim2 + 1), npix = dim1 * dim2;
  double *tth = (double *)malloc(npix * sizeof(double));
  double *chi = (double *)malloc(npix * sizeof(double));
  double *rad = (double *)malloc(npix * sizeof(double));
  double *tc = (double *)malloc(nc * sizeof(double));
  double *ac = (double *)malloc(nc * sizeof(double));
  double *rc = (double *)malloc(nc * sizeof(double));
  or_centres(p, dim1, dim2, tth, chi);
  or_convert_unit(unit, p->wavelength, tth, npix, rad);
  or_corners(p, dim1, dim2, tc, ac);
  or_convert_unit(unit, p->wavelength, tc, nc, rc);
  double r_lo = rlo, r_hi = rhi;
  if (!has_rr) pyfai_range(rc, nc, &r_lo, &r_hi);
  double a_lo = has_ar ? alo : -OR_PI, a_hi = has_ar ? ahi : OR_PI;
  grid2d g;
  if (grid_new(&g, n_radial, n_azimuthal, r_lo, r_hi, a_lo, a_hi)) {
    free(tth); free(chi); free(rad); free(tc); free(ac); free(rc);
    return 2;
  }
  g.clamp_index_rounding = 1;
  if (edges) { edges[0] = g.r0; edges[1] = g.r1; edges[2] = g.a0; edges[3] = g.a1; }
  trig t = make_trig(p);
  size_t nb = n_radial * n_azimuthal;
  bucket *bk = (bucket *)calloc(nb, sizeof(bucket));
  size_t w = dim2 + 1;
  for (size_t pix = 0; pix < npix; pix++) {
    size_t b0 = pix / dim2 * w + pix % dim2;
    size_t idx[4] = {b0, b0 + 1, b0 + w + 1, b0 + w};
    int axis_on_edge = 0, zero_r = 0;
    double winding = 0.0;
    for (int i = 0; i < 4; i++) {
      double delta = ac[idx[(i + 1) % 4]] - ac[idx[i]];
      delta = rem_euclid(delta + OR_PI, OR_TAU) - OR_PI;
      if (fabs(delta) >= OR_PI - 64.0 * OR_EPS) axis_on_edge = 1;
      winding += delta;
      if (rc[idx[i]] == 0.0) zero_r = 1;
    }
    if (fabs(winding) > OR_PI || axis_on_edge || zero_r) {
      sample_pixel_2d(p, &t, dim2, pix, unit, &g, bk);
      continue;
    }
    size_t pidx[4] = {b0, b0 + w, b0 + w + 1, b0 + 1}; /* pyFAI order */
    double pp[4][2];
    for (int i = 0; i < 4; i++) {
      pp[i][0] = rc[pidx[i]];
      pp[i][1] = ac[pidx[i]];
    }
    if (pyfai_split2d(pp, &g, pix, bk)) continue;
    double center = g.a0 + rem_euclid(chi[pix] - g.a0, OR_TAU);
    double poly[4][2];
    for (int i = 0; i < 4; i++) {
      poly[i][0] = rc[idx[i]];
      poly[i][1] = center + rem_euclid(ac[idx[i]] - center + OR_PI, OR_TAU) - OR_PI;
    }
    double area = fabs(polygon_area((const double(*)[2])poly, 4));
    if (area > 0.0) {
      split_polygon_2d((const double(*)[2])poly, area, &g, pix, bk);
    } else {
      long long bb = grid_bin(&g, rad[pix], chi[pix]);
      if (bb >= 0) bucket_push(&bk[bb], pix, 1.0);
    }
  }
  buckets_to_lut(bk, nb, 1, g.r0, grid_rstep(&g), out);
  free(bk);
  free(tth); free(chi); free(rad); free(tc); free(ac); free(rc);
  return 0;
}

/* --------------------------------------------------- execution and gather */

/* prepared.rs:163-193: entries of masked pixels dropped, order kept. An
 * all-zero mask keeps the table as is (copied here). */
void or_fixed_table(const or_lut *in, const uint8_t *mask, or_lut *out) {
  out->nbins = in->nbins;
  out->hist_min = in->hist_min;
  out->hist_step = in->hist_step;
  out->row_ptr = (uint64_t *)malloc((in->nbins + 1) * sizeof(uint64_t));
  out->col = (uint64_t *)malloc((in->nnz ? in->nnz : 1) * sizeof(uint64_t));
  out->w = in->w ? (double *)malloc((in->nnz ? in->nnz : 1) * sizeof(double)) : NULL;
  size_t k = 0;
  out->row_ptr[0] = 0;
  for (size_t r = 0; r < in->nbins; r++) {
    for (uint64_t e = in->row_ptr[r]; e < in->row_ptr[r + 1]; e++) {
      uint64_t j = in->col[e];
      if (mask && mask[j]) continue;
      out->col[k] = j;
      if (in->w) out->w[k] = in->w[e];
      k++;
    }
    out->row_ptr[r + 1] = k;
  }
  out->nnz = k;
}

/* prepared.rs:229-262 with row<FAST=false, NORM> (1011-1121): sequential
 * n = n + w * norm[j] in row order (n + norm[j] for unweighted tables). */
void or_denominators(const or_lut *lut, const double *norm, double *den) {
  for (size_t r = 0; r < lut->nbins; r++) {
    double n = 0.0;
    for (uint64_t e = lut->row_ptr[r]; e < lut->row_ptr[r + 1]; e++) {
      uint64_t j = lut->col[e];
      if (lut->w) n = n + lut->w[e] * norm[j];
      else n = n + norm[j];
    }
    den[r] = n;
  }
}

/* prepared.rs:985-997 (normalize_bin) and 1011-1121 (strict row). */
void or_gather(const or_lut *lut, const double *den, const double *signal, int with_var,
               double *intensity, double *sigma) {
  for (size_t r = 0; r < lut->nbins; r++) {
    double s = 0.0, v = 0.0;
    for (uint64_t e = lut->row_ptr[r]; e < lut->row_ptr[r + 1]; e++) {
      double x = signal[lut->col[e]];
      if (lut->w) {
        double w = lut->w[e];
        s = s + w * x;
        if (with_var) v = v + w * w * fmax(x, 0.0);
      } else {
        s += x;
        if (with_var) v += fmax(x, 0.0);
      }
    }
    double n = den[r];
    intensity[r] = n == 0.0 ? 0.0 : s / n;
    if (with_var && sigma) {
      double vv = n == 0.0 ? 0.0 : v / (n * n);
      sigma[r] = sqrt(fmax(vv, 0.0));
    }
  }
}

/* azimuthal.rs:48-118. Invalid contributions are skipped; origin is the
 * first valid corrected value; empty rows give (0, 0, NaN). */
void or_azimuthal_std(const or_lut *lut, const uint8_t *mask, const double *norm,
                      const double *signal, int errors, double *intensity, double *sigma,
                      double *std_out) {
  for (size_t r = 0; r < lut->nbins; r++) {
    int found = 0;
    double origin = 0.0, sum_signal = 0.0, sum_weight = 0.0, sum_shifted = 0.0, sum_var = 0.0;
    for (int pass = 0; pass < 2; pass++) {
      for (uint64_t e = lut->row_ptr[r]; e < lut->row_ptr[r + 1]; e++) {
        uint64_t j = lut->col[e];
        if (mask && mask[j]) continue;
        double n = norm[j];
        double c = lut->w ? lut->w[e] : 1.0;
        double s = signal[j];
        double x = s / n;
        double w = c * n;
        if (!(isfinite(s) && isfinite(n) && n > 0.0 && isfinite(c) && c > 0.0 && isfinite(x) &&
              isfinite(w) && w > 0.0))
          continue;
        if (pass == 0) {
          origin = x;
          found = 1;
          break;
        }
        sum_signal += c * s;
        sum_weight += w;
        sum_shifted += w * (x - origin);
        if (errors) sum_var += c * c * fmax(s, 0.0);
      }
      if (!found) break;
    }
    if (!found) {
      intensity[r] = 0.0;
      if (sigma) sigma[r] = 0.0;
      std_out[r] = NAN;
      continue;
    }
    double mean_shifted = sum_shifted / sum_weight;
    double m2 = 0.0;
    for (uint64_t e = lut->row_ptr[r]; e < lut->row_ptr[r + 1]; e++) {
      uint64_t j = lut->col[e];
      if (mask && mask[j]) continue;
      double n = norm[j];
      double c = lut->w ? lut->w[e] : 1.0;
      double s = signal[j];
      double x = s / n;
      double w = c * n;
      if (!(isfinite(s) && isfinite(n) && n > 0.0 && isfinite(c) && c > 0.0 && isfinite(x) &&
            isfinite(w) && w > 0.0))
        continue;
      double delta = (x - origin) - mean_shifted;
      m2 += w * delta * delta;
    }
    intensity[r] = sum_signal / sum_weight;
    if (sigma) sigma[r] = sqrt(fmax(sum_var, 0.0)) / sum_weight;
    std_out[r] = sqrt(fmax(m2 / sum_weight, 0.0));
  }
}

double or_wrap_tau(double x) { return wrap_tau(x); }

/* Exposes split1d (pyfai_1d.rs:259-354) on one polygon for known-answer
 * checks; returns the entry count written to bins/weights (capacity cap). */
int or_pyfai_split1d_sector(const double *poly8, double r0, double r1, size_t nbins, double a0,
                            double a1, int sector, double sign, uint64_t *bins, double *weights,
                            size_t cap) {
  double poly[4][2];
  for (int i = 0; i < 4; i++) {
    poly[i][0] = poly8[2 * i];
    poly[i][1] = poly8[2 * i + 1];
  }
  bucket *bk = (bucket *)calloc(nbins, sizeof(bucket));
  split1d(poly, r0, r1, nbins, a0, a1, sector, sign, 0, bk);
  size_t n = 0;
  for (size_t b = 0; b < nbins; b++) {
    for (size_t e = 0; e < bk[b].len; e++)
      if (n < cap) {
        bins[n] = b;
        weights[n] = bk[b].w[e];
        n++;
      }
    free(bk[b].col);
    free(bk[b].w);
  }
  free(bk);
  return (int)n;
}

int or_pyfai_split1d(const double *poly8, double r0, double r1, size_t nbins, uint64_t *bins,
                     double *weights, size_t cap) {
  return or_pyfai_split1d_sector(poly8, r0, r1, nbins, -OR_PI, OR_PI, 0, 1.0, bins, weights,
                                 cap);
}

/* grid.rs:26-222 IntegrationGrid2d::new + bin for known-answer checks: flags
 * bit 0 clamped_radial_end, 1 clamped_index_rounding, 2 clamped_radial_outside,
 * 3 clamped_azimuthal_outside. Returns the bin, -1 for none, -2 for an
 * invalid grid. */
long long or_grid2d_bin(size_t nr, size_t na, double r0, double r1, double a0, double a1,
                        int flags, double radial, double az) {
  grid2d g;
  if (grid_new(&g, nr, na, r0, r1, a0, a1)) return -2;
  g.clamp_radial_end = (flags & 1) != 0;
  g.clamp_index_rounding = (flags & 2) != 0;
  g.clamp_radial_outside = (flags & 4) != 0;
  g.clamp_azimuthal_outside = (flags & 8) != 0;
  return grid_bin(&g, radial, az);
}
