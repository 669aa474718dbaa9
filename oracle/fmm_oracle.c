/*
 * fmm_oracle.c -- CPU restatement of the reference laplace-fmm path.
 *
 * TEST INFRASTRUCTURE ONLY (see fmm_oracle.h).  Every function cites the
 * reference file:line whose algorithm it follows (paths relative to the
 * reference repository root).  Arithmetic is written in the reference's
 * evaluation order and compiled with -ffp-contract=off so that rounding
 * follows the reference's scalar code; thread scheduling (TBB there, OpenMP
 * here) only changes the order of floating-point reductions.
 */
#include "fmm_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

typedef double complex cplx;

#define PI_D 3.14159265358979323846
#define FORCE_EPS 1.e-4                 /* src/common/defs.h:24 */
#define FORCE_EPS2 (FORCE_EPS * FORCE_EPS) /* src/common/defs.h:25 */

static inline double ni(int n) { return (n & 1) ? -1.0 : 1.0; } /* simple_math.h:96-99 */
static inline int tri(int n) { return n * (n + 1) / 2; }

/* ===================================================================== */
/* constant tables                                                       */
/* ===================================================================== */

/* constexpr Newton square root, src/common/special_functions.h:46-54:
 * iterate c <- (c + x/c)/2 from c = x until |(prev-c)/c| < 1e-16. */
static double newton_sqrt(double x)
{
	double prev = 0.0, curr = x;
	for (int it = 0; it < 400; ++it) {
		double rel = (prev - curr) / curr;
		if (rel < 1.e-16 && rel > -1.e-16) return curr;
		double next = 0.5 * (curr + x / curr);
		prev = curr;
		curr = next;
	}
	return curr;
}

#define KMAX ORC_MAX_TERMS_3D
static double g_knm[KMAX * (KMAX + 1) / 2];           /* knm_wrapper<18>, special_functions.h:56-85 */
static double g_anm[2 * KMAX * (2 * KMAX + 1) / 2];   /* anm_wrapper<18>, :87-116 */
static double g_wig_aux[100 * 101 / 2];               /* compute_wignerd_coef(100), special_functions.cpp:194-207 */
static int g_scalar_ready = 0; /* Knm/Anm/aux coefficient tables built */
static int g_tables_ready = 0; /* plus the direction-keyed maps */

static void build_scalar_tables(void)
{
	/* Knm = sqrt((n-m)!/(n+m)!) by the product recurrence, then Newton sqrt. */
	g_knm[0] = 1.0;
	for (int n = 1; n < KMAX; ++n) {
		for (int m = 0; m <= n - 1; ++m)
			g_knm[tri(n) + m] = g_knm[tri(n - 1) + m] * (n - m) / (n + m);
		g_knm[tri(n) + n] = g_knm[tri(n) + n - 1] / (2 * n);
	}
	for (int n = 1; n < KMAX; ++n)
		for (int m = 0; m <= n; ++m)
			g_knm[tri(n) + m] = newton_sqrt(g_knm[tri(n) + m]);

	/* Anm = (-1)^n / sqrt((n-m)!(n+m)!) for n < 36. */
	g_anm[0] = 1.0;
	for (int n = 1; n < 2 * KMAX; ++n) {
		for (int m = 0; m <= n - 1; ++m)
			g_anm[tri(n) + m] = g_anm[tri(n - 1) + m] * (n - m) * (n + m);
		g_anm[tri(n) + n] = g_anm[tri(n) + n - 1] * 2 * n;
	}
	for (int n = 1; n < 2 * KMAX; ++n)
		for (int m = 0; m <= n; ++m)
			g_anm[tri(n) + m] = ni(n) / newton_sqrt(g_anm[tri(n) + m]);

	g_wig_aux[0] = 1;
	for (int n = 1; n < 100; ++n) {
		g_wig_aux[tri(n)] = sqrt((2 * n - 1.) / (2 * n)) * g_wig_aux[tri(n - 1)];
		for (int m = 1; m <= n; ++m)
			g_wig_aux[tri(n) + m] = sqrt((n - m + 1.) / (n + m)) * g_wig_aux[tri(n) + m - 1];
	}
}

static inline double Knm(int n, int m) { return g_knm[tri(n) + m]; }
static inline double Anm(int n, int m) { return g_anm[tri(n) + m]; }

/* m2lcoef_wrapper, special_functions.h:118-138: A(n,k) A(j,k) (-1)^(n+k) / A(j+n,0). */
double orc_m2lcoef(int j, int k, int n)
{
	if (!g_scalar_ready) orc_init();
	return Anm(n, k) * Anm(j, k) * ni(n + k) / Anm(j + n, 0);
}

void orc_knm(int N, double* out)
{
	if (!g_scalar_ready) orc_init();
	for (int i = 0; i < tri(N); ++i) out[i] = g_knm[i];
}

void orc_anm(int N, double* out)
{
	if (!g_scalar_ready) orc_init();
	for (int i = 0; i < tri(2 * N); ++i) out[i] = g_anm[i];
}

/* ComputePnm, special_functions.h:12-31: associated Legendre P_n^m(cos theta)
 * with the Condon-Shortley sign on the diagonal, n < N. */
void orc_pnm(int N, double theta, double* P)
{
	double x = cos(theta);
	double y = sin(theta);
	P[0] = 1;
	for (int n = 1; n < N; ++n) {
		for (int m = 0; m <= n - 2; ++m)
			P[tri(n) + m] = ((2 * n - 1) * x * P[tri(n - 1) + m] - (n + m - 1) * P[tri(n - 2) + m]) / (n - m);
		P[tri(n) + n - 1] = P[tri(n - 1) + n - 1] * x * (2 * n - 1);
		P[tri(n) + n] = P[tri(n - 1) + n - 1] * y * (1 - 2 * n);
	}
}

/* DecToSph, simple_math.h:101-117: r = |v| + 1e-12, theta = acos(z/r),
 * phi = atan2(y,x) with the x = 0 special cases.  The radius offset is the
 * reference's (g_r_eps = 1e-12) unless orc_set_r_eps changes it. */
static double g_r_eps = 1e-12;
void orc_set_r_eps(double e) { g_r_eps = e; }
double orc_get_r_eps(void) { return g_r_eps; }

void orc_dec_to_sph(const double* v, double* c)
{
	const double eps = 1e-12;
	c[0] = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + g_r_eps;
	c[1] = acos(v[2] / c[0]);
	if (fabs(v[0]) + fabs(v[1]) < eps)
		c[2] = 0;
	else if (fabs(v[0]) < eps)
		c[2] = v[1] / fabs(v[1]) * PI_D * 0.5;
	else
		c[2] = atan2(v[1], v[0]);
}

/* doublehash, utils.h:120-123 */
int orc_doublehash(double v) { return (int)(v * 100); }

static inline int wig_base(int n) { return n * (1 + n) * (4 * n - 1) / 6; }

/* ComputeWignerD, special_functions.cpp:211-351.  d^n_{m,k}(theta) for
 * 0<=m<=n<N, -n<=k<=n at wig_base(n) + m(2n+1) + (k+n), including the
 * eps(k) sign and the theta<0 / theta>pi/2 symmetries of the reference. */
void orc_wignerd(int N, double theta, double* D)
{
	if (!g_scalar_ready) orc_init();
	const int size = N * (1 + N) * (4 * N - 1) / 6;
	memset(D, 0, sizeof(double) * size);
	int isnegative = theta < 0;
	theta = fabs(theta);
	int closetopi = 0;
	if (theta > PI_D / 2) {
		theta = PI_D - theta;
		closetopi = 1;
	}
	if (fabs(theta) < 1e-3) {
		for (int n = 0; n < N; ++n) {
			int i0 = wig_base(n) + n;
			for (int m = 0; m <= n; ++m) D[i0 + (2 * n + 1) * m + m] = 1.0;
		}
	} else {
		const double* g = g_wig_aux;
		double s = sin(theta);
		double c = 1.0 + cos(theta);
		double f = s / c;
		double* sn = (double*)calloc(2 * N, sizeof(double));
		double* cn = (double*)calloc(2 * N, sizeof(double));
		sn[N] = cn[N] = 1.0;
		for (int i = 1; i < N; ++i) {
			sn[N + i] = sn[N + i - 1] * s;
			sn[N - i] = sn[N - i + 1] / s;
			cn[N + i] = cn[N + i - 1] * c;
			cn[N - i] = cn[N - i + 1] / c;
		}
		/* 1) column k = n */
		for (int n = 0; n < N; ++n) {
			int i0 = wig_base(n) + 2 * n;
			for (int m = 0; m <= n; ++m)
				D[i0 + m * (2 * n + 1)] = ni(n + m) * g[tri(n) + m] * cn[N + m] * sn[N + n - m];
		}
		/* 2) row m = n, k = n-1 .. -n */
		for (int n = 0; n < N; ++n) {
			int i0 = wig_base(n) + n * (2 * n + 1) + n;
			for (int m = n; m > -n; --m)
				D[i0 + m - 1] = (n + m) / sqrt(n * (n + 1.) - m * (m - 1.)) * f * D[i0 + m];
		}
		/* 3) rows m = n-1 .. 0 */
		for (int n = 0; n < N; ++n) {
			int i0 = wig_base(n) + n;
			int i1 = 2 * n + 1;
			for (int m = n - 1; m >= 0; --m) {
				int i2 = m * i1, i3 = (m + 1) * i1;
				for (int k = n; k > -n; --k)
					D[i0 + i2 + k - 1] = sqrt((n * (n + 1.) - m * (m + 1.)) / (n * (n + 1.) - k * (k - 1.))) * D[i0 + i3 + k] +
						(m + k) / sqrt(n * (n + 1.) - k * (k - 1.)) * f * D[i0 + i2 + k];
			}
		}
		free(sn);
		free(cn);
	}
	if (closetopi) {
		double* Dc = (double*)malloc(sizeof(double) * size);
		memcpy(Dc, D, sizeof(double) * size);
		for (int n = 1; n < N; ++n) {
			int i0 = wig_base(n), i1 = 2 * n + 1;
			for (int m = 0; m <= n; ++m) {
				int i2 = i0 + i1 * m + n;
				for (int k = -n; k <= n; ++k) D[i2 + k] = ni(n + m) * Dc[i2 - k];
			}
		}
		free(Dc);
	}
	for (int n = 1; n < N; ++n) {
		int i0 = wig_base(n), i1 = 2 * n + 1;
		for (int m = 0; m <= n; ++m) {
			int i2 = i0 + i1 * m + n;
			for (int k = -n; k <= n; ++k) {
				int kk = k < 0 ? -k : k;
				D[i2 + k] *= ni((k - kk) / 2);
				if (isnegative) D[i2 + k] *= ni(m - k);
			}
		}
	}
}

/* ---- direction tables keyed by doublehash, special_functions.cpp:363-381 ---- */
#define HKEY_OFF 700
#define HKEY_SPAN 1401
#define WIG_SIZE_MAX (KMAX * (1 + KMAX) * (4 * KMAX - 1) / 6)
static double* g_dmat[HKEY_SPAN];
static cplx* g_rexp[HKEY_SPAN];

static void compute_exp(int N, double phi, cplx* out)        /* ComputeExp, special_functions.cpp:353-359 */
{
	for (int i = 0; i < N; ++i) out[i] = CMPLX(cos(i * phi), sin(i * phi));
}

static void put_dmat(int key, double theta)
{
	int idx = key + HKEY_OFF;
	if (!g_dmat[idx]) g_dmat[idx] = (double*)malloc(sizeof(double) * WIG_SIZE_MAX);
	orc_wignerd(KMAX, theta, g_dmat[idx]);
}

static void put_rexp(int key, double phi)
{
	int idx = key + HKEY_OFF;
	if (!g_rexp[idx]) g_rexp[idx] = (cplx*)malloc(sizeof(cplx) * KMAX);
	compute_exp(KMAX, phi, g_rexp[idx]);
}

void orc_init(void)
{
	if (g_tables_ready) return;
	#pragma omp critical(orc_init_lock)
	{
		if (!g_tables_ready) {
			if (!g_scalar_ready) { build_scalar_tables(); g_scalar_ready = 1; }
			for (int i = -3; i < 4; ++i)
				for (int j = -3; j < 4; ++j)
					for (int k = -3; k < 4; ++k) {
						if (abs(i) <= 1 && abs(j) <= 1 && abs(k) <= 1) continue;
						double v[3] = {i, j, k}, rho[3];
						orc_dec_to_sph(v, rho);
						double theta = rho[1], phi = rho[2];
						int key = orc_doublehash(theta);
						put_dmat(key, theta);
						put_dmat(-key, -theta);
						key = orc_doublehash(phi);
						put_rexp(key, phi);
						put_rexp(-key, -phi);
					}
			g_tables_ready = 1;
		}
	}
}

static const double* lookup_dmat(double theta)
{
	int key = orc_doublehash(theta);
	if (key < -HKEY_OFF || key > HKEY_OFF) return NULL;
	return g_dmat[key + HKEY_OFF];
}

static const cplx* lookup_rexp(double phi)
{
	int key = orc_doublehash(phi);
	if (key < -HKEY_OFF || key > HKEY_OFF) return NULL;
	return g_rexp[key + HKEY_OFF];
}

int orc_dmatrix_lookup(double theta, int N, double* out)
{
	orc_init();
	const double* d = lookup_dmat(theta);
	if (!d) return -1;
	memcpy(out, d, sizeof(double) * (N * (1 + N) * (4 * N - 1) / 6));
	return 0;
}

/* ===================================================================== */
/* 3D operators, src/cpu/multipole3d.cpp                                  */
/* ===================================================================== */

/* RotateZ accumulate form, multipole3d.cpp:65-77: b += a * e^{i m phi}. */
static void rotate_z_acc(int N, const cplx* a, const cplx* e, cplx* b)
{
	for (int n = 0; n < N; ++n)
		for (int m = 0; m <= n; ++m) b[tri(n) + m] += a[tri(n) + m] * e[m];
}

/* RotateY, multipole3d.cpp:79-113. */
static void rotate_y(int N, const cplx* a, const double* d, cplx* res)
{
	for (int n = 0; n < N; ++n) {
		int i0 = (n * (5 + n * (3 + 4 * n))) / 6;
		int i1 = 2 * n + 1;
		for (int m = 0; m <= n; ++m) {
			int i2 = i0 + i1 * m;
			double vr = 0.0, vi = 0.0;
			for (int k = 1; k <= n; ++k) {
				double z1 = d[i2 + k], z2 = d[i2 - k];
				cplx w = a[tri(n) + k];
				vr = creal(w) * (z1 + z2) + vr;
				vi = cimag(w) * (z1 - z2) + vi;
			}
			cplx a0 = a[tri(n)];
			vr += d[i2] * creal(a0);
			vi += d[i2] * cimag(a0);
			res[tri(n) + m] = CMPLX(vr, vi);
		}
	}
}

enum { OP_M2M = 0, OP_M2L = 1, OP_L2L = 2 };

/* M2M / M2L / L2L share the rotate - axial translate - rotate back frame:
 * multipole3d.cpp:115-150 (M2M), :152-190 (M2L), :192-226 (L2L). */
static int rtr_op(int kind, int N, const cplx* a, const double* p0, cplx* b, cplx* w1, cplx* w2)
{
	double rho[3];
	orc_dec_to_sph(p0, rho);
	const double theta = rho[1], phi = rho[2];
	const double* dm1 = lookup_dmat(theta);
	const double* dm2 = lookup_dmat(-theta);
	const cplx* e1 = lookup_rexp(phi);
	const cplx* e2 = lookup_rexp(-phi);
	if (!dm1 || !dm2 || !e1 || !e2) return -1; /* the reference throws std::out_of_range */
	const int Nx2 = tri(N);
	memset(w2, 0, sizeof(cplx) * Nx2);
	rotate_z_acc(N, a, e1, w2);
	rotate_y(N, w2, dm1, w1);
	memset(w2, 0, sizeof(cplx) * Nx2);
	double rn[2 * ORC_MAX_TERMS_3D + 1];
	if (kind == OP_M2L) {
		rn[0] = 1;
		for (int i = 1; i < 2 * N + 1; ++i) rn[i] = rn[i - 1] / rho[0];
		for (int j = 0; j < N; ++j)
			for (int k = 0; k <= j; ++k)
				for (int n = k; n < N; ++n)
					w2[tri(j) + k] += (orc_m2lcoef(j, k, n) * rn[j + n + 1]) * w1[tri(n) + k];
	} else {
		rn[0] = 1;
		for (int i = 1; i < N; ++i) rn[i] = rn[i - 1] * rho[0];
		if (kind == OP_M2M) {
			for (int j = 0; j < N; ++j)
				for (int k = 0; k <= j; ++k)
					for (int n = 0; j - n >= k; ++n) {
						int jn = j - n;
						w2[tri(j) + k] += Anm(n, 0) * Anm(jn, k) * rn[n] / Anm(j, k) * w1[tri(jn) + k];
					}
		} else {
			for (int j = 0; j < N; ++j)
				for (int k = 0; k <= j; ++k)
					for (int n = j; n < N; ++n)
						w2[tri(j) + k] += Anm(n - j, 0) * Anm(j, k) * ni(n + j) * rn[n - j] / Anm(n, k) * w1[tri(n) + k];
		}
	}
	rotate_y(N, w2, dm2, w1);
	rotate_z_acc(N, w1, e2, b);
	return 0;
}

/* Multipole (P2M), multipole3d.cpp:20-49: M_nm += q r^n K_nm P_n^m e^{-im phi}. */
static void p2m_particles(int N, int64_t n_p, const double* xyzq_sorted_x, const double* y, const double* z,
                          const double* q, const double* c, cplx* a, double* P, cplx* eim)
{
	for (int64_t i = 0; i < n_p; ++i) {
		double v[3] = {xyzq_sorted_x[i] - c[0], y[i] - c[1], z[i] - c[2]}, rho[3];
		orc_dec_to_sph(v, rho);
		double r = rho[0], rn = r;
		orc_pnm(N, rho[1], P);
		for (int m = 0; m < N; ++m) {
			double mphi = m * rho[2];
			eim[m] = CMPLX(cos(mphi), -sin(mphi));
		}
		for (int n = 1; n < N; ++n) {
			for (int m = 0; m <= n; ++m)
				a[tri(n) + m] += q[i] * rn * Knm(n, m) * P[tri(n) + m] * eim[m];
			rn *= r;
		}
		a[0] += q[i];
	}
}

/* L2P of ComputeLeaves<double>, multipole3d.cpp:1152-1208 (potential and
 * -grad via the spherical derivatives, dividing by sin(theta)). */
static void l2p_one(int N, const cplx* inner, const double* c, const double* p, double* P, cplx* eim,
                    double* pot, double* f)
{
	double v[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]}, dr[3];
	orc_dec_to_sph(v, dr);
	orc_pnm(N + 1, dr[1], P);
	for (int m = 0; m <= N; ++m) {
		double x = m * dr[2];
		eim[m] = CMPLX(cos(x), sin(x));
	}
	double x = cos(dr[1]);
	double y = sin(dr[1]);
	double rn = 1. / dr[0];
	double phi = 0.0, t0 = 0, t1 = 0, t2 = 0;
	for (int n = 0; n < N; ++n) {
		for (int m = 0; m <= n; ++m) {
			cplx Ynm = P[tri(n) + m] * eim[m] * Knm(n, m);
			cplx coef = inner[tri(n) + m] * rn;
			double pv = creal(dr[0] * Ynm * coef);
			phi += pv;
			if (m != 0) phi += pv;
			cplx dth = Knm(n, m) * eim[m] * ((1 - m + n) * P[tri(n + 1) + m] - (n + 1) * x * P[tri(n) + m]) / y;
			cplx dph = ((double)m * I) * Ynm;
			double a0 = creal((double)n * coef * Ynm), a1 = creal(coef * dth), a2 = creal(coef * dph);
			t0 -= a0; t1 -= a1; t2 -= a2;
			if (m != 0) { t0 -= a0; t1 -= a1; t2 -= a2; }
		}
		rn *= dr[0];
	}
	*pot += phi;
	double s1 = sin(dr[1]), c2 = cos(dr[2]), c1 = cos(dr[1]), s2 = sin(dr[2]);
	f[0] += t0 * s1 * c2 + t1 * c1 * c2 - t2 * s2 / s1;
	f[1] += t0 * s1 * s2 + t1 * c1 * s2 + t2 * c2 / s1;
	f[2] += t0 * c1 - t1 * s1;
}

void orc_p2m3d(int N, int64_t n, const double* xyz, const double* q, const double* centre, double* M)
{
	orc_init();
	double* x = (double*)malloc(sizeof(double) * (3 * n + 1));
	double *y = x + n, *z = x + 2 * n;
	for (int64_t i = 0; i < n; ++i) { x[i] = xyz[3 * i]; y[i] = xyz[3 * i + 1]; z[i] = xyz[3 * i + 2]; }
	double P[(ORC_MAX_TERMS_3D + 2) * (ORC_MAX_TERMS_3D + 3) / 2];
	cplx eim[ORC_MAX_TERMS_3D + 2];
	p2m_particles(N, n, x, y, z, q, centre, (cplx*)M, P, eim);
	free(x);
}

static int op_public(int kind, int N, const double* a, const double* p0, double* b)
{
	orc_init();
	cplx w1[ORC_MAX_TERMS_3D * (ORC_MAX_TERMS_3D + 1) / 2], w2[ORC_MAX_TERMS_3D * (ORC_MAX_TERMS_3D + 1) / 2];
	return rtr_op(kind, N, (const cplx*)a, p0, (cplx*)b, w1, w2);
}
void orc_m2m3d(int N, const double* a, const double* p0, double* b) { op_public(OP_M2M, N, a, p0, b); }
void orc_m2l3d(int N, const double* a, const double* p0, double* b) { op_public(OP_M2L, N, a, p0, b); }
void orc_l2l3d(int N, const double* a, const double* p0, double* b) { op_public(OP_L2L, N, a, p0, b); }

void orc_l2p3d(int N, const double* L, const double* centre, int64_t n, const double* xyz, double* pot, double* field)
{
	orc_init();
	double P[(ORC_MAX_TERMS_3D + 2) * (ORC_MAX_TERMS_3D + 3) / 2];
	cplx eim[ORC_MAX_TERMS_3D + 2];
	for (int64_t i = 0; i < n; ++i) l2p_one(N, (const cplx*)L, centre, xyz + 3 * i, P, eim, pot + i, field + 3 * i);
}

/* ===================================================================== */
/* Morton tree, src/cpu/morton_tree.cpp (dim = 2 or 3)                    */
/* ===================================================================== */

typedef struct {
	int64_t size;
	uint64_t* key;
	int64_t* sb; int64_t* se;   /* source_range */
	int64_t* tb; int64_t* te;   /* target_range (leaves) */
	int64_t* parent;
	double* centre;             /* dim per cell */
	int64_t* near_off; int64_t* near_idx;
	int64_t* far_off; int64_t* far_idx;
} orc_level;

struct orc_tree3d {
	int dim, depth;
	int64_t n, targets_num;
	double shift[3], scale;
	uint64_t* key;      /* sorted particle keys (after partition) */
	int64_t* pos_map;   /* positions_map: original -> sorted position */
	double* pos;        /* sorted positions, dim per particle */
	double* q;          /* sorted charges */
	orc_level lv[ORC_MAX_DEPTH];
};

/* ExpandBits<2>/<3>, morton_tree.cpp:48-82: floor(2^depth x) with the bits
 * spread by the magic-mask ladder. */
static uint64_t expand_bits(int dim, double x, int depth)
{
	uint64_t bit_shift = (1ull << depth);
	uint64_t v = (uint64_t)((double)bit_shift * x);
	if (dim == 2) {
		v = ((v << 16) | v) & 0xffff0000ffffull;
		v = ((v << 8) | v) & 0xff00ff00ff00ffull;
		v = ((v << 4) | v) & 0xf0f0f0f0f0f0f0full;
		v = ((v << 2) | v) & 0x3333333333333333ull;
		v = ((v << 1) | v) & 0x5555555555555555ull;
	} else {
		v = ((v << 32) | v) & 0xffff00000000ffffull;
		v = ((v << 16) | v) & 0xff0000ff0000ffull;
		v = ((v << 8) | v) & 0xf00f00f00f00f00full;
		v = ((v << 4) | v) & 0x30c30c30c30c30c3ull;
		v = ((v << 2) | v) & 0x9249249249249249ull;
	}
	return v;
}

/* MortonCode, morton_tree.cpp:84-99: x is the high bit of each triple. */
static uint64_t morton_code(int dim, const double* p, int depth, const double* shift, double scale)
{
	if (dim == 2) {
		double x = (p[0] - shift[0]) / scale, y = (p[1] - shift[1]) / scale;
		return (expand_bits(2, x, depth) << 1) | expand_bits(2, y, depth);
	}
	double x = (p[0] - shift[0]) / scale, y = (p[1] - shift[1]) / scale, z = (p[2] - shift[2]) / scale;
	return (expand_bits(3, x, depth) << 2) | (expand_bits(3, y, depth) << 1) | expand_bits(3, z, depth);
}

/* PointFromMortonCode, morton_tree.cpp:101-145: dyadic cell centre. */
static void point_from_code(int dim, uint64_t val, int depth, const double* shift, double scale, double* out)
{
	double step = 1.0 / (1 << (depth + 1));
	double c[3] = {step, step, step};
	step *= 2;
	for (int i = 0; i < depth; ++i) {
		for (int a = dim - 1; a >= 0; --a) {
			c[a] += step * (double)(val & 1);
			val >>= 1;
		}
		step *= 2;
	}
	for (int a = 0; a < dim; ++a) {
		double t = c[a] * scale;
		out[a] = t + shift[a];
	}
}

static double sqdist(int dim, const double* a, const double* b)
{
	double s = 0.0;
	for (int i = 0; i < dim; ++i) { double d = a[i] - b[i]; s = (i == 0) ? d * d : s + d * d; }
	return s;
}

typedef struct { int64_t cnt, cap; int64_t* v; } ivec;
static void ivec_push(ivec* a, int64_t x)
{
	if (a->cnt == a->cap) { a->cap = a->cap ? 2 * a->cap : 32; a->v = (int64_t*)realloc(a->v, sizeof(int64_t) * a->cap); }
	a->v[a->cnt++] = x;
}

/* FillNeighbours, morton_tree.cpp:147-208, plus the half-list dedupe of
 * :474-489 when every particle is a source. */
static void fill_neighbours(orc_tree3d* t)
{
	const int dim = t->dim;
	t->lv[0].near_off = (int64_t*)calloc(t->lv[0].size + 1, sizeof(int64_t));
	t->lv[0].near_idx = (int64_t*)malloc(sizeof(int64_t));
	t->lv[0].far_off = (int64_t*)calloc(t->lv[0].size + 1, sizeof(int64_t));
	t->lv[0].far_idx = (int64_t*)malloc(sizeof(int64_t));
	if (t->depth > 1) {
		orc_level* L1 = &t->lv[1];
		int64_t s = L1->size;
		L1->near_off = (int64_t*)calloc(s + 1, sizeof(int64_t));
		L1->near_idx = (int64_t*)malloc(sizeof(int64_t) * (s * (s - 1) + 1));
		L1->far_off = (int64_t*)calloc(s + 1, sizeof(int64_t));
		L1->far_idx = (int64_t*)malloc(sizeof(int64_t));
		int64_t c = 0;
		for (int64_t i = 0; i < s; ++i) {
			for (int64_t k = 0; k < s; ++k) if (k != i) L1->near_idx[c++] = k;
			L1->near_off[i + 1] = c;
		}
	}
	for (int i = 2; i < t->depth; ++i) {
		double dw = sqrt(dim + 1.e-5) * t->scale / (1 << i);
		dw *= dw;
		orc_level* P = &t->lv[i - 1];
		orc_level* C = &t->lv[i];
		ivec* nl = (ivec*)calloc(C->size, sizeof(ivec));
		ivec* fl = (ivec*)calloc(C->size, sizeof(ivec));
		#pragma omp parallel for schedule(dynamic, 64)
		for (int64_t j = 0; j < C->size; ++j) {
			const double* cc = C->centre + dim * j;
			int64_t par = C->parent[j];
			for (int64_t a = P->near_off[par]; a < P->near_off[par + 1]; ++a) {
				int64_t pn = P->near_idx[a];
				for (int64_t k = P->sb[pn]; k < P->se[pn]; ++k) {
					if (sqdist(dim, C->centre + dim * k, cc) < dw) ivec_push(&nl[j], k);
					else ivec_push(&fl[j], k);
				}
			}
			for (int64_t k = P->sb[par]; k < P->se[par]; ++k) {
				if (k == j) continue;
				if (sqdist(dim, C->centre + dim * k, cc) < dw) ivec_push(&nl[j], k);
				else ivec_push(&fl[j], k);
			}
		}
		C->near_off = (int64_t*)calloc(C->size + 1, sizeof(int64_t));
		C->far_off = (int64_t*)calloc(C->size + 1, sizeof(int64_t));
		for (int64_t j = 0; j < C->size; ++j) {
			C->near_off[j + 1] = C->near_off[j] + nl[j].cnt;
			C->far_off[j + 1] = C->far_off[j] + fl[j].cnt;
		}
		C->near_idx = (int64_t*)malloc(sizeof(int64_t) * (C->near_off[C->size] + 1));
		C->far_idx = (int64_t*)malloc(sizeof(int64_t) * (C->far_off[C->size] + 1));
		for (int64_t j = 0; j < C->size; ++j) {
			if (nl[j].cnt) memcpy(C->near_idx + C->near_off[j], nl[j].v, sizeof(int64_t) * nl[j].cnt);
			if (fl[j].cnt) memcpy(C->far_idx + C->far_off[j], fl[j].v, sizeof(int64_t) * fl[j].cnt);
			free(nl[j].v); free(fl[j].v);
		}
		free(nl); free(fl);
	}
	if (t->targets_num == 0 && t->depth > 1) {
		/* Sequential erase loop of the reference keeps, for leaf i, exactly the
		 * near neighbours with a larger index, in their original order. */
		orc_level* L = &t->lv[t->depth - 1];
		int64_t c = 0;
		int64_t* off = (int64_t*)calloc(L->size + 1, sizeof(int64_t));
		for (int64_t j = 0; j < L->size; ++j) {
			for (int64_t a = L->near_off[j]; a < L->near_off[j + 1]; ++a)
				if (L->near_idx[a] > j) L->near_idx[c++] = L->near_idx[a];
			off[j + 1] = c;
		}
		free(L->near_off);
		L->near_off = off;
	}
}

typedef struct { uint64_t k; int64_t i; } kv_pair;
static int kv_cmp(const void* a, const void* b)
{
	const kv_pair* x = (const kv_pair*)a;
	const kv_pair* y = (const kv_pair*)b;
	if (x->k != y->k) return x->k < y->k ? -1 : 1;
	return x->i < y->i ? -1 : (x->i > y->i);
}

/* MortonTree constructor, morton_tree.cpp:286-493. */
static orc_tree3d* tree_build(int dim, int64_t n, const double* pos_in, const double* q_in, int depth)
{
	if (n <= 0 || depth < 1 || depth > ORC_MAX_DEPTH) return NULL;
	orc_tree3d* t = (orc_tree3d*)calloc(1, sizeof(orc_tree3d));
	t->dim = dim; t->depth = depth; t->n = n;
	/* box_extent, :218-284 and :293-301 */
	double mn[3], mx[3];
	for (int a = 0; a < dim; ++a) mn[a] = mx[a] = pos_in[a];
	for (int64_t i = 0; i < n; ++i)
		for (int a = 0; a < dim; ++a) {
			double v = pos_in[dim * i + a];
			if (v < mn[a]) mn[a] = v;
			if (v > mx[a]) mx[a] = v;
		}
	double dw = mx[0] - mn[0];
	for (int a = 1; a < dim; ++a) if (mx[a] - mn[a] > dw) dw = mx[a] - mn[a];
	for (int a = 0; a < dim; ++a) t->shift[a] = mn[a] - 1.e-7;
	t->scale = dw + 1.e-6;

	/* keys at depth-1, :316-319 */
	uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * n);
	#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < n; ++i) key[i] = morton_code(dim, pos_in + dim * i, depth - 1, t->shift, t->scale);

	/* stable counting sort, :329-358, in its serial execution order (the
	 * reference's atomic bucket cursors make the in-leaf order of its
	 * threaded run timing dependent; the stable order is the deterministic
	 * one).  Deep trees, whose dense counter array the reference would also
	 * allocate, use an equivalent stable (key, index) comparison sort. */
	t->pos_map = (int64_t*)malloc(sizeof(int64_t) * n);
	t->key = (uint64_t*)malloc(sizeof(uint64_t) * n);
	t->pos = (double*)malloc(sizeof(double) * dim * n);
	t->q = (double*)malloc(sizeof(double) * n);
	if (dim * (depth - 1) <= 24) {
		int64_t nb = (int64_t)1 << (dim * (depth - 1));
		int64_t* cnt = (int64_t*)calloc(nb + 1, sizeof(int64_t));
		for (int64_t i = 0; i < n; ++i) cnt[1 + key[i]]++;
		for (int64_t i = 1; i <= nb; ++i) cnt[i] += cnt[i - 1];
		for (int64_t i = 0; i < n; ++i) t->pos_map[i] = cnt[key[i]]++;
		free(cnt);
	} else {
		kv_pair* kv = (kv_pair*)malloc(sizeof(kv_pair) * n);
		for (int64_t i = 0; i < n; ++i) { kv[i].k = key[i]; kv[i].i = i; }
		qsort(kv, n, sizeof(kv_pair), kv_cmp);
		for (int64_t p = 0; p < n; ++p) t->pos_map[kv[p].i] = p;
		free(kv);
	}
	for (int64_t i = 0; i < n; ++i) {
		int64_t p = t->pos_map[i];
		t->key[p] = key[i];
		for (int a = 0; a < dim; ++a) t->pos[dim * p + a] = pos_in[dim * i + a];
		t->q[p] = q_in[i];
	}
	free(key);

	/* leaf level, :364-393 */
	for (int l = 0; l < depth; ++l) {
		int64_t cap = ((int64_t)1 << (dim * l));
		if (cap > n || l > 20) cap = n;
		orc_level* L = &t->lv[l];
		L->key = (uint64_t*)malloc(sizeof(uint64_t) * cap);
		L->sb = (int64_t*)malloc(sizeof(int64_t) * cap);
		L->se = (int64_t*)malloc(sizeof(int64_t) * cap);
		L->parent = (int64_t*)calloc(cap, sizeof(int64_t));
		L->centre = (double*)malloc(sizeof(double) * dim * cap);
	}
	{
		orc_level* L = &t->lv[depth - 1];
		int64_t c = 0;
		for (int64_t j = 0; j < n; ++j) {
			if (j == 0 || t->key[j] != t->key[j - 1]) {
				if (j > 0) { L->se[c - 1] = j; }
				L->key[c] = t->key[j];
				L->sb[c] = j;
				++c;
			}
		}
		L->se[c - 1] = n;
		L->size = c;
		for (int64_t j = 0; j < c; ++j) point_from_code(dim, L->key[j], depth - 1, t->shift, t->scale, L->centre + dim * j);
	}
	/* upper levels, :395-423 */
	for (int i = depth - 2; i >= 0; --i) {
		orc_level* T = &t->lv[i];
		orc_level* B = &t->lv[i + 1];
		int64_t c = 0;
		for (int64_t j = 0; j < B->size; ++j) {
			uint64_t pk = B->key[j] >> dim;
			if (j == 0 || pk != (B->key[j - 1] >> dim)) {
				if (j > 0) T->se[c - 1] = j;
				T->key[c] = pk;
				T->sb[c] = j;
				++c;
			}
			B->parent[j] = c - 1;
		}
		T->se[c - 1] = B->size;
		T->size = c;
		for (int64_t j = 0; j < c; ++j) point_from_code(dim, T->key[j], i, t->shift, t->scale, T->centre + dim * j);
	}
	/* source/target partition, :426-468: inside each leaf, q != 0 first. */
	{
		orc_level* L = &t->lv[depth - 1];
		int64_t* inv = (int64_t*)malloc(sizeof(int64_t) * n);
		for (int64_t i = 0; i < n; ++i) inv[t->pos_map[i]] = i;
		L->tb = (int64_t*)malloc(sizeof(int64_t) * L->size);
		L->te = (int64_t*)malloc(sizeof(int64_t) * L->size);
		int64_t tn = 0;
		for (int64_t k = 0; k < L->size; ++k) {
			int64_t s0 = L->sb[k], s1 = L->se[k], tb = s0;
			for (int64_t i = s0; i < s1; ++i) {
				if (t->q[i] != 0.0) {
					for (int a = 0; a < dim; ++a) {
						double tmp = t->pos[dim * tb + a]; t->pos[dim * tb + a] = t->pos[dim * i + a]; t->pos[dim * i + a] = tmp;
					}
					double tq = t->q[tb]; t->q[tb] = t->q[i]; t->q[i] = tq;
					int64_t ti = inv[tb]; inv[tb] = inv[i]; inv[i] = ti;
					++tb;
				}
			}
			L->tb[k] = tb; L->te[k] = s1;
			L->se[k] = tb;
			tn += s1 - tb;
		}
		for (int64_t i = 0; i < n; ++i) t->pos_map[inv[i]] = i;
		free(inv);
		t->targets_num = tn;
	}
	fill_neighbours(t);
	return t;
}

orc_tree3d* orc_tree3d_build(int64_t n, const double* xyz, const double* q, int depth)
{
	return tree_build(3, n, xyz, q, depth);
}

/* MortonTree<2> (src/cpu/morton_tree.cpp with dim = 2); read back through the
 * orc_tree3d_* accessors, which follow the tree's dimension. */
orc_tree3d* orc_tree2d_build(int64_t n, const double* xy, const double* q, int depth)
{
	return tree_build(2, n, xy, q, depth);
}

void orc_tree3d_free(orc_tree3d* t)
{
	if (!t) return;
	for (int l = 0; l < t->depth; ++l) {
		orc_level* L = &t->lv[l];
		free(L->key); free(L->sb); free(L->se); free(L->tb); free(L->te); free(L->parent); free(L->centre);
		free(L->near_off); free(L->near_idx); free(L->far_off); free(L->far_idx);
	}
	free(t->key); free(t->pos_map); free(t->pos); free(t->q);
	free(t);
}

int orc_tree3d_depth(const orc_tree3d* t) { return t->depth; }

void orc_tree3d_info(const orc_tree3d* t, int64_t* level_sizes, double* shift_scale, int64_t* targets_num)
{
	for (int l = 0; l < t->depth; ++l) level_sizes[l] = t->lv[l].size;
	for (int a = 0; a < t->dim; ++a) shift_scale[a] = t->shift[a];
	shift_scale[t->dim] = t->scale;
	*targets_num = t->targets_num;
}

void orc_tree3d_particles(const orc_tree3d* t, uint64_t* keys, int64_t* pos_map, double* xyzq)
{
	const int d = t->dim;
	for (int64_t i = 0; i < t->n; ++i) {
		if (keys) keys[i] = t->key[i];
		if (pos_map) pos_map[i] = t->pos_map[i];
		if (xyzq) {
			for (int a = 0; a < d; ++a) xyzq[(d + 1) * i + a] = t->pos[d * i + a];
			xyzq[(d + 1) * i + d] = t->q[i];
		}
	}
}

void orc_tree3d_level(const orc_tree3d* t, int level, uint64_t* keys, int64_t* src_range,
                      int64_t* tgt_range, int64_t* parent, double* centres)
{
	const orc_level* L = &t->lv[level];
	for (int64_t j = 0; j < L->size; ++j) {
		if (keys) keys[j] = L->key[j];
		if (src_range) { src_range[2 * j] = L->sb[j]; src_range[2 * j + 1] = L->se[j]; }
		if (tgt_range && L->tb) { tgt_range[2 * j] = L->tb[j]; tgt_range[2 * j + 1] = L->te[j]; }
		if (parent) parent[j] = L->parent[j];
		if (centres) for (int a = 0; a < t->dim; ++a) centres[t->dim * j + a] = L->centre[t->dim * j + a];
	}
}

int64_t orc_tree3d_list_size(const orc_tree3d* t, int level, int far)
{
	const orc_level* L = &t->lv[level];
	const int64_t* off = far ? L->far_off : L->near_off;
	return off ? off[L->size] : 0;
}

void orc_tree3d_lists(const orc_tree3d* t, int level, int far, int64_t* offsets, int64_t* idx)
{
	const orc_level* L = &t->lv[level];
	const int64_t* off = far ? L->far_off : L->near_off;
	const int64_t* v = far ? L->far_idx : L->near_idx;
	if (!off) { for (int64_t j = 0; j <= L->size; ++j) offsets[j] = 0; return; }
	for (int64_t j = 0; j <= L->size; ++j) offsets[j] = off[j];
	for (int64_t a = 0; a < off[L->size]; ++a) idx[a] = v[a];
}

int orc_num_threads(void) { return omp_get_max_threads(); }
/* Thread count of the oracle's parallel loops; one thread makes every sum's
 * order fixed (the P2P half lists and the level loops then run in index
 * order), so results repeat bit for bit. */
void orc_set_num_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }

/* ===================================================================== */
/* 3D solve, multipole3d.cpp:1296-1374                                    */
/* ===================================================================== */

static int auto_terms_3d(double eps, int N)
{
	if (N == ORC_AUTO) {
		eps = eps > 1.e-11 ? eps : 1.e-11;
		int a = (int)(54.2 - 10.73 * sqrt(log(2.4 * eps) + 25.5));
		return a < ORC_MAX_TERMS_3D ? a : ORC_MAX_TERMS_3D;
	}
	return N < ORC_MAX_TERMS_3D ? N : ORC_MAX_TERMS_3D;
}

/* The 3D solve.  pterms (NULL: one order N for every box) gives per-box
 * truncation orders, a restatement of the GPU's per-box orders (no
 * reference counterpart): every leaf takes the largest request of its
 * particles (pterms in input order, one per output), a box with children the
 * largest of its children, and the M2L and L2L into a box of level >= 4 run
 * with the box's own order (levels 2-3 and P2M, M2M, L2P with the
 * largest order; coefficients beyond a box's order stay zero). */
static int fmm3d_impl(int64_t n_src, const double* src_xyz, const double* src_q,
                      int64_t n_tgt, const double* tgt_xyz,
                      double eps, int N_, int depth_, double* pot_out, double* field_out, double* timings,
                      const int32_t* pterms)
{
	orc_init();
	double T0 = omp_get_wtime(), t;
	int N = auto_terms_3d(eps, N_);
	const int64_t n = n_src + n_tgt;
	const int64_t out_n = n_tgt ? n_tgt : n_src;
	int depth = depth_;
	if (depth_ == ORC_AUTO) {
		double d = 1 + log((double)n / 100.0) / log(8.0);
		depth = (int)(d > 2.0 ? d : 2.0);
	}
	if (N < 1 || depth < 1 || depth > 12 || n <= 0) return -1;
	/* particle array: targets (q = 0) first, then sources, :1311-1319 */
	double* pos = (double*)malloc(sizeof(double) * 3 * n);
	double* q = (double*)malloc(sizeof(double) * n);
	for (int64_t i = 0; i < n_tgt; ++i) { memcpy(pos + 3 * i, tgt_xyz + 3 * i, 24); q[i] = 0.0; }
	memcpy(pos + 3 * n_tgt, src_xyz, sizeof(double) * 3 * n_src);
	memcpy(q + n_tgt, src_q, sizeof(double) * n_src);
	orc_tree3d* tr = tree_build(3, n, pos, q, depth);
	free(pos); free(q);
	if (!tr) return -1;
	double t_tree = omp_get_wtime() - T0;
	/* per-box orders: bt[level][box] */
	int** bt = NULL;
	if (pterms) {
		bt = (int**)calloc(depth, sizeof(int*));
		for (int l = 0; l < depth; ++l) bt[l] = (int*)calloc(tr->lv[l].size ? tr->lv[l].size : 1, sizeof(int));
		orc_level* LL = &tr->lv[depth - 1];
		int64_t* leaf_of = (int64_t*)malloc(sizeof(int64_t) * n);
		for (int64_t li = 0; li < LL->size; ++li) {
			for (int64_t s = LL->sb[li]; s < LL->se[li]; ++s) leaf_of[s] = li;
			for (int64_t s = LL->tb[li]; s < LL->te[li]; ++s) leaf_of[s] = li;
		}
		N = 1;
		for (int64_t i = 0; i < out_n; ++i) {
			int v = pterms[i] < 1 ? 1 : (pterms[i] > ORC_MAX_TERMS_3D ? ORC_MAX_TERMS_3D : pterms[i]);
			int64_t li = leaf_of[tr->pos_map[i]];
			if (v > bt[depth - 1][li]) bt[depth - 1][li] = v;
			if (v > N) N = v;
		}
		free(leaf_of);
		for (int l = depth - 2; l >= 0; --l)
			for (int64_t j = 0; j < tr->lv[l].size; ++j)
				for (int64_t k = tr->lv[l].sb[j]; k < tr->lv[l].se[j]; ++k)
					if (bt[l + 1][k] > bt[l][j]) bt[l][j] = bt[l + 1][k];
	}
	const int Nx2 = tri(N);
#define BOX_TERMS(i, j) ((bt && (i) >= 4) ? bt[i][j] : N)

	cplx** outer = (cplx**)calloc(depth, sizeof(cplx*));
	cplx** inner = (cplx**)calloc(depth, sizeof(cplx*));
	for (int l = 0; l < depth; ++l) {
		outer[l] = (cplx*)calloc((size_t)tr->lv[l].size * Nx2, sizeof(cplx));
		inner[l] = (cplx*)calloc((size_t)tr->lv[l].size * Nx2, sizeof(cplx));
	}
	int err = 0;
	const double* X = tr->pos;
	/* Upward, :228-318 */
	t = omp_get_wtime();
	{
		orc_level* L = &tr->lv[depth - 1];
		#pragma omp parallel
		{
			double P[(ORC_MAX_TERMS_3D + 2) * (ORC_MAX_TERMS_3D + 3) / 2];
			cplx eim[ORC_MAX_TERMS_3D + 2];
			double* buf = (double*)malloc(sizeof(double) * 4 * 4096);
			#pragma omp for schedule(dynamic, 16)
			for (int64_t i = 0; i < L->size; ++i) {
				int64_t s0 = L->sb[i], cntp = L->se[i] - s0;
				double* bx = cntp <= 4096 ? buf : (double*)malloc(sizeof(double) * 4 * cntp);
				for (int64_t k = 0; k < cntp; ++k) {
					bx[k] = X[3 * (s0 + k)]; bx[cntp + k] = X[3 * (s0 + k) + 1];
					bx[2 * cntp + k] = X[3 * (s0 + k) + 2]; bx[3 * cntp + k] = tr->q[s0 + k];
				}
				p2m_particles(N, cntp, bx, bx + cntp, bx + 2 * cntp, bx + 3 * cntp, L->centre + 3 * i, outer[depth - 1] + Nx2 * i, P, eim);
				if (bx != buf) free(bx);
			}
			free(buf);
		}
	}
	double t_p2m = omp_get_wtime() - t;
	t = omp_get_wtime();
	for (int i = depth - 2; i >= 0; --i) {
		orc_level* T = &tr->lv[i];
		orc_level* B = &tr->lv[i + 1];
		#pragma omp parallel
		{
			cplx w1[171], w2[171];
			#pragma omp for schedule(dynamic, 16) reduction(|:err)
			for (int64_t j = 0; j < T->size; ++j)
				for (int64_t k = T->sb[j]; k < T->se[j]; ++k) {
					double p0[3] = {B->centre[3 * k] - T->centre[3 * j], B->centre[3 * k + 1] - T->centre[3 * j + 1],
					                B->centre[3 * k + 2] - T->centre[3 * j + 2]};
					err |= rtr_op(OP_M2M, N, outer[i + 1] + k * Nx2, p0, outer[i] + j * Nx2, w1, w2);
				}
		}
	}
	double t_m2m = omp_get_wtime() - t;
	/* Downward, :320-412: M2L for levels 2.., then L2L for levels 3.. */
	t = omp_get_wtime();
	for (int i = 2; i < depth; ++i) {
		orc_level* L = &tr->lv[i];
		#pragma omp parallel
		{
			cplx w1[171], w2[171];
			#pragma omp for schedule(dynamic, 16) reduction(|:err)
			for (int64_t j = 0; j < L->size; ++j)
				for (int64_t a = L->far_off[j]; a < L->far_off[j + 1]; ++a) {
					int64_t s = L->far_idx[a];
					double p0[3] = {L->centre[3 * s] - L->centre[3 * j], L->centre[3 * s + 1] - L->centre[3 * j + 1],
					                L->centre[3 * s + 2] - L->centre[3 * j + 2]};
					err |= rtr_op(OP_M2L, BOX_TERMS(i, j), outer[i] + s * Nx2, p0, inner[i] + Nx2 * j, w1, w2);
				}
		}
	}
	for (int i = 3; i < depth; ++i) {
		orc_level* T = &tr->lv[i - 1];
		orc_level* L = &tr->lv[i];
		#pragma omp parallel
		{
			cplx w1[171], w2[171];
			#pragma omp for schedule(dynamic, 16) reduction(|:err)
			for (int64_t j = 0; j < L->size; ++j) {
				int64_t p = L->parent[j];
				double p0[3] = {T->centre[3 * p] - L->centre[3 * j], T->centre[3 * p + 1] - L->centre[3 * j + 1],
				                T->centre[3 * p + 2] - L->centre[3 * j + 2]};
				err |= rtr_op(OP_L2L, BOX_TERMS(i, j), inner[i - 1] + Nx2 * p, p0, inner[i] + Nx2 * j, w1, w2);
			}
		}
	}
	double t_down = omp_get_wtime() - t;

	/* ComputeLeaves<double>, :855-1294 */
	t = omp_get_wtime();
	orc_level* L = &tr->lv[depth - 1];
	const int use_src = tr->targets_num == 0;
	const int nth = omp_get_max_threads();
	double* bufs = (double*)calloc((size_t)nth * 4 * n, sizeof(double)); /* per-thread pot, fx, fy, fz */
	#pragma omp parallel
	{
		const int tid = omp_get_thread_num();
		double* bp = bufs + (size_t)tid * 4 * n;
		double* bf = bp + n; /* 3 per particle */
		double P[(ORC_MAX_TERMS_3D + 2) * (ORC_MAX_TERMS_3D + 3) / 2];
		cplx eim[ORC_MAX_TERMS_3D + 2];
		#pragma omp for schedule(dynamic, 8)
		for (int64_t li = 0; li < L->size; ++li) {
			int64_t p1b = use_src ? L->sb[li] : L->tb[li];
			int64_t p1e = use_src ? L->se[li] : L->te[li];
			for (int64_t j = p1b; j < p1e; ++j) {
				const double xt = X[3 * j], yt = X[3 * j + 1], zt = X[3 * j + 2], qt = tr->q[j];
				double fx = 0, fy = 0, fz = 0, pt = 0;
				/* the pair numerics of avx_inv_dr, avx.h:128-146 */
				#define ORC_PAIR(K, MUTUAL)                                                        \
					do {                                                                             \
						const int64_t k_ = (K);                                                      \
						double dx = xt - X[3 * k_], dy = yt - X[3 * k_ + 1], dz = zt - X[3 * k_ + 2];  \
						double r2 = dx * dx + dy * dy;                                               \
						r2 = r2 + dz * dz;                                                           \
						r2 = r2 > FORCE_EPS2 ? r2 : FORCE_EPS2;                                      \
						double i2 = 1.0 / r2, i1 = sqrt(i2);                                         \
						double qi = tr->q[k_] * i1, qi3 = i2 * qi;                                   \
						fx = dx * qi3 + fx; fy = dy * qi3 + fy; fz = dz * qi3 + fz; pt = qi + pt;    \
						if (MUTUAL) {                                                                \
							double qj = qt * i1, qj3 = i2 * qj;                                      \
							bf[3 * k_] -= dx * qj3; bf[3 * k_ + 1] -= dy * qj3; bf[3 * k_ + 2] -= dz * qj3; \
							bp[k_] += qj;                                                            \
						}                                                                            \
					} while (0)
				if (use_src) {
					ORC_PAIR(j, 0); /* self term, clamped */
					for (int64_t k = j + 1; k < p1e; ++k) ORC_PAIR(k, 1);
					for (int64_t a = L->near_off[li]; a < L->near_off[li + 1]; ++a) {
						int64_t nb = L->near_idx[a];
						for (int64_t k = L->sb[nb]; k < L->se[nb]; ++k) ORC_PAIR(k, 1);
					}
				} else {
					for (int64_t k = L->sb[li]; k < L->se[li]; ++k) ORC_PAIR(k, 0);
					for (int64_t a = L->near_off[li]; a < L->near_off[li + 1]; ++a) {
						int64_t nb = L->near_idx[a];
						for (int64_t k = L->sb[nb]; k < L->se[nb]; ++k) ORC_PAIR(k, 0);
					}
				}
				#undef ORC_PAIR
				bf[3 * j] += fx; bf[3 * j + 1] += fy; bf[3 * j + 2] += fz;
				bp[j] += pt;
			}
			const cplx* in = inner[depth - 1] + li * Nx2;
			for (int64_t j = p1b; j < p1e; ++j) l2p_one(N, in, L->centre + 3 * li, X + 3 * j, P, eim, bp + j, bf + 3 * j);
		}
	}
	/* reduce per-thread buffers, :1212-1281 */
	#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < n; ++i)
		for (int th = 1; th < nth; ++th) {
			const double* o = bufs + (size_t)th * 4 * n;
			bufs[i] += o[i];
			bufs[n + 3 * i] += o[n + 3 * i];
			bufs[n + 3 * i + 1] += o[n + 3 * i + 1];
			bufs[n + 3 * i + 2] += o[n + 3 * i + 2];
		}
	/* un-permute the first out_n original particles, :1288-1293 */
	for (int64_t i = 0; i < out_n; ++i) {
		int64_t p = tr->pos_map[i];
		if (pot_out) pot_out[i] = bufs[p];
		if (field_out) { field_out[3 * i] = bufs[n + 3 * p]; field_out[3 * i + 1] = bufs[n + 3 * p + 1]; field_out[3 * i + 2] = bufs[n + 3 * p + 2]; }
	}
	double t_leaf = omp_get_wtime() - t;
	free(bufs);
	for (int l = 0; l < depth; ++l) { free(outer[l]); free(inner[l]); }
	free(outer); free(inner);
	if (bt) {
		for (int l = 0; l < depth; ++l) free(bt[l]);
		free(bt);
	}
#undef BOX_TERMS
	orc_tree3d_free(tr);
	if (timings) {
		timings[0] = t_tree; timings[1] = t_p2m; timings[2] = t_m2m; timings[3] = t_down; timings[4] = t_leaf;
		timings[5] = omp_get_wtime() - T0;
	}
	return err ? -2 : 0;
}

int orc_fmm3d(int64_t n_src, const double* src_xyz, const double* src_q,
              int64_t n_tgt, const double* tgt_xyz,
              double eps, int N_, int depth_, double* pot_out, double* field_out, double* timings)
{
	return fmm3d_impl(n_src, src_xyz, src_q, n_tgt, tgt_xyz, eps, N_, depth_, pot_out, field_out, timings, NULL);
}

int orc_fmm3d_orders(int64_t n_src, const double* src_xyz, const double* src_q, int64_t n_tgt, const double* tgt_xyz,
                     const int32_t* pterms, int depth, double* pot_out, double* field_out)
{
	return fmm3d_impl(n_src, src_xyz, src_q, n_tgt, tgt_xyz, 1e-8, 1, depth, pot_out, field_out, NULL, pterms);
}

/* ===================================================================== */
/* 2D solve, src/cpu/multipole.cpp                                        */
/* ===================================================================== */

static double binom60[60 * 60]; /* BinomNewton_wrapper<60>, simple_math.h:25-49 */
static int binom_ready = 0;
static void build_binom(void)
{
	if (binom_ready) return;
	binom60[0] = 1.0;
	for (int i = 1; i < 60; ++i) {
		binom60[i * 60] = 1.0;
		binom60[i * 60 + i] = 1.0;
		for (int j = 1; j < i; ++j) binom60[i * 60 + j] = binom60[(i - 1) * 60 + j] + binom60[(i - 1) * 60 + j - 1];
	}
	binom_ready = 1;
}
static inline double binom(int p, int q) { return binom60[p * 60 + q]; }

static cplx cpow_int(cplx base, unsigned e) /* MyPow, simple_math.h:81-94 */
{
	cplx r = 1;
	while (e) { if (e & 1) r *= base; e >>= 1; base *= base; }
	return r;
}
static cplx cdiv_plain(cplx a, cplx b) /* divComp, multipole.cpp:55-64 */
{
	double ax = creal(a), ay = cimag(a), bx = creal(b), by = cimag(b);
	double zn = bx * bx + by * by;
	return CMPLX((ax * bx + ay * by) / zn, (ay * bx - ax * by) / zn);
}

int orc_fmm2d(int64_t n_src, const double* src_xy, const double* src_q,
              int64_t n_tgt, const double* tgt_xy,
              double eps, int N_, int depth_, double* pot_out, double* field_out, double* timings)
{
	build_binom();
	double T0 = omp_get_wtime();
	int N = N_;
	if (N_ == ORC_AUTO) N = (int)(-1.18 * log(1.65 * eps)); /* multipole.cpp:478-481 */
	if (N > ORC_MAX_TERMS_2D) N = ORC_MAX_TERMS_2D;
	const int64_t n = n_src + n_tgt, out_n = n_tgt ? n_tgt : n_src;
	int depth = depth_;
	if (depth_ == ORC_AUTO) {
		double d = 2 + log((double)n / 100.0) / log(4.0);
		depth = (int)(d > 3.0 ? d : 3.0);
	}
	if (N < 1 || depth < 1 || depth > 16 || n <= 0) return -1;
	double* pos = (double*)malloc(sizeof(double) * 2 * n);
	double* q = (double*)malloc(sizeof(double) * n);
	for (int64_t i = 0; i < n_tgt; ++i) { pos[2 * i] = tgt_xy[2 * i]; pos[2 * i + 1] = tgt_xy[2 * i + 1]; q[i] = 0.0; }
	memcpy(pos + 2 * n_tgt, src_xy, sizeof(double) * 2 * n_src);
	memcpy(q + n_tgt, src_q, sizeof(double) * n_src);
	orc_tree3d* tr = tree_build(2, n, pos, q, depth);
	free(pos); free(q);
	if (!tr) return -1;
	double t_tree = omp_get_wtime() - T0, t = omp_get_wtime();
	cplx** outer = (cplx**)calloc(depth, sizeof(cplx*));
	cplx** inner = (cplx**)calloc(depth, sizeof(cplx*));
	for (int l = 0; l < depth; ++l) {
		outer[l] = (cplx*)calloc((size_t)tr->lv[l].size * N, sizeof(cplx));
		inner[l] = (cplx*)calloc((size_t)tr->lv[l].size * N, sizeof(cplx));
	}
	const double* X = tr->pos;
	/* Multipole, multipole.cpp:22-36 */
	{
		orc_level* L = &tr->lv[depth - 1];
		#pragma omp parallel for schedule(dynamic, 16)
		for (int64_t i = 0; i < L->size; ++i) {
			cplx z0 = CMPLX(L->centre[2 * i], L->centre[2 * i + 1]);
			cplx* a = outer[depth - 1] + N * i;
			for (int k = 1; k < N; ++k) {
				cplx s = 0;
				for (int64_t p = L->sb[i]; p < L->se[i]; ++p) s = s + tr->q[p] * cpow_int(CMPLX(X[2 * p], X[2 * p + 1]) - z0, k);
				a[k] = -s / (double)k;
			}
			cplx s = 0;
			for (int64_t p = L->sb[i]; p < L->se[i]; ++p) s = s + tr->q[p];
			a[0] = s;
		}
	}
	double t_p2m = omp_get_wtime() - t;
	t = omp_get_wtime();
	/* M2M, multipole.cpp:38-53 */
	for (int i = depth - 2; i >= 0; --i) {
		orc_level* T = &tr->lv[i];
		orc_level* B = &tr->lv[i + 1];
		#pragma omp parallel
		{
			cplx w[ORC_MAX_TERMS_2D];
			#pragma omp for schedule(dynamic, 16)
			for (int64_t j = 0; j < T->size; ++j)
				for (int64_t k = T->sb[j]; k < T->se[j]; ++k) {
					cplx z0 = CMPLX(B->centre[2 * k] - T->centre[2 * j], B->centre[2 * k + 1] - T->centre[2 * j + 1]);
					const cplx* a = outer[i + 1] + N * k;
					cplx* b = outer[i] + N * j;
					w[0] = 1.0;
					for (int s = 1; s < N; ++s) w[s] = w[s - 1] * z0;
					for (int jj = 1; jj < N; ++jj) {
						for (int s = 1; s <= jj; ++s) b[jj] += a[s] * w[jj - s] * binom(jj - 1, s - 1);
						b[jj] -= a[0] * w[jj] / (double)jj;
					}
					b[0] += a[0];
				}
		}
	}
	double t_m2m = omp_get_wtime() - t;
	t = omp_get_wtime();
	/* Downward, multipole.cpp:179-231: per level L2L (assign) then M2L. */
	for (int i = 2; i < depth; ++i) {
		orc_level* T = &tr->lv[i - 1];
		orc_level* L = &tr->lv[i];
		#pragma omp parallel
		{
			cplx w[ORC_MAX_TERMS_2D];
			#pragma omp for schedule(dynamic, 16)
			for (int64_t j = 0; j < L->size; ++j) {
				int64_t p = L->parent[j];
				cplx* b = inner[i] + N * j;
				{ /* L2L, :86-97 */
					cplx z0 = CMPLX(T->centre[2 * p] - L->centre[2 * j], T->centre[2 * p + 1] - L->centre[2 * j + 1]);
					const cplx* a = inner[i - 1] + N * p;
					for (int s = 0; s < N; ++s) b[s] = a[s];
					for (int jj = 0; jj < N; ++jj)
						for (int k = N - jj - 1; k < N - 1; ++k) b[k] -= z0 * b[k + 1];
				}
				for (int64_t f = L->far_off[j]; f < L->far_off[j + 1]; ++f) { /* M2L, :66-84 */
					int64_t s = L->far_idx[f];
					cplx z0 = CMPLX(L->centre[2 * s] - L->centre[2 * j], L->centre[2 * s + 1] - L->centre[2 * j + 1]);
					const cplx* a = outer[i] + N * s;
					w[0] = a[0];
					b[0] += a[0] * conj(clog(-z0));
					for (int k = 1; k < N; ++k) {
						w[k] = ni(k) * a[k] / cpow_int(z0, k);
						b[0] += w[k];
					}
					for (int k = 1; k < N; ++k) {
						cplx tmp = 0.0;
						for (int jj = 1; jj < N; ++jj) tmp += w[jj] * binom(k + jj - 1, jj - 1);
						b[k] += cdiv_plain(-a[0] / (double)k + tmp, cpow_int(z0, k));
					}
				}
			}
		}
	}
	double t_down = omp_get_wtime() - t;
	t = omp_get_wtime();
	/* ComputePotentials / ComputeForces, multipole.cpp:233-449 */
	orc_level* L = &tr->lv[depth - 1];
	const int use_src = tr->targets_num == 0;
	const int nth = omp_get_max_threads();
	double* bufs = (double*)calloc((size_t)nth * 3 * n, sizeof(double)); /* pot, fre, fim */
	#pragma omp parallel
	{
		const int tid = omp_get_thread_num();
		double* bp = bufs + (size_t)tid * 3 * n;
		double* bf = bp + n;
		#pragma omp for schedule(dynamic, 8)
		for (int64_t li = 0; li < L->size; ++li) {
			int64_t p1b = use_src ? L->sb[li] : L->tb[li];
			int64_t p1e = use_src ? L->se[li] : L->te[li];
			for (int64_t j = p1b; j < p1e; ++j) {
				const double xt = X[2 * j], yt = X[2 * j + 1], qt = tr->q[j];
				double ph = 0;
				cplx fo = 0;
				#define ORC_PAIR2(K, MUTUAL)                                                  \
					do {                                                                        \
						const int64_t k_ = (K);                                                 \
						cplx dz = CMPLX(xt - X[2 * k_], yt - X[2 * k_ + 1]);                    \
						double nz = creal(dz) * creal(dz) + cimag(dz) * cimag(dz);              \
						double mz = nz > FORCE_EPS2 ? nz : FORCE_EPS2;                          \
						double lg = 0.5 * log(mz);                                               \
						cplx inv = conj(dz) / mz;                                                \
						ph += tr->q[k_] * lg;                                                    \
						fo += tr->q[k_] * inv;                                                   \
						if (MUTUAL) { bp[k_] += qt * lg; bf[2 * k_] -= creal(qt * inv); bf[2 * k_ + 1] -= cimag(qt * inv); } \
					} while (0)
				if (use_src) {
					ph = qt * 0.5 * log(FORCE_EPS2);
					for (int64_t k = j + 1; k < p1e; ++k) ORC_PAIR2(k, 1);
					for (int64_t a = L->near_off[li]; a < L->near_off[li + 1]; ++a) {
						int64_t nb = L->near_idx[a];
						for (int64_t k = L->sb[nb]; k < L->se[nb]; ++k) ORC_PAIR2(k, 1);
					}
				} else {
					for (int64_t k = L->sb[li]; k < L->se[li]; ++k) ORC_PAIR2(k, 0);
					for (int64_t a = L->near_off[li]; a < L->near_off[li + 1]; ++a) {
						int64_t nb = L->near_idx[a];
						for (int64_t k = L->sb[nb]; k < L->se[nb]; ++k) ORC_PAIR2(k, 0);
					}
				}
				#undef ORC_PAIR2
				bp[j] += ph;
				bf[2 * j] += creal(fo); bf[2 * j + 1] += cimag(fo);
			}
			const cplx* in = inner[depth - 1] + li * N;
			cplx zc = CMPLX(L->centre[2 * li], L->centre[2 * li + 1]);
			for (int64_t j = p1b; j < p1e; ++j) {
				cplx dz = CMPLX(X[2 * j], X[2 * j + 1]) - zc;
				cplx zp = 1.0;
				double ph = 0.0;
				for (int k = 0; k < N; ++k) { ph += creal(in[k]) * creal(zp) - cimag(in[k]) * cimag(zp); zp *= dz; }
				bp[j] += ph;
				cplx fo = 0.0;
				zp = 1.0;
				for (int k = 1; k < N; ++k) { fo += (double)k * in[k] * zp; zp *= dz; }
				bf[2 * j] += creal(fo); bf[2 * j + 1] += cimag(fo);
			}
		}
	}
	#pragma omp parallel for schedule(static)
	for (int64_t i = 0; i < n; ++i)
		for (int th = 1; th < nth; ++th) {
			const double* o = bufs + (size_t)th * 3 * n;
			bufs[i] += o[i];
			bufs[n + 2 * i] += o[n + 2 * i];
			bufs[n + 2 * i + 1] += o[n + 2 * i + 1];
		}
	for (int64_t i = 0; i < out_n; ++i) {
		int64_t p = tr->pos_map[i];
		if (pot_out) pot_out[i] = bufs[p];
		if (field_out) { field_out[2 * i] = bufs[n + 2 * p]; field_out[2 * i + 1] = bufs[n + 2 * p + 1]; }
	}
	double t_leaf = omp_get_wtime() - t;
	free(bufs);
	for (int l = 0; l < depth; ++l) { free(outer[l]); free(inner[l]); }
	free(outer); free(inner);
	orc_tree3d_free(tr);
	if (timings) {
		timings[0] = t_tree; timings[1] = t_p2m; timings[2] = t_m2m; timings[3] = t_down; timings[4] = t_leaf;
		timings[5] = omp_get_wtime() - T0;
	}
	return 0;
}

/* ===================================================================== */
/* direct sums and error statistics                                      */
/* ===================================================================== */

/* Potential3d / Force3d, simple_math.h:130-139; ComputeExact :218-281. */
void orc_direct3d(int64_t n_src, const double* S, const double* Q,
                  int64_t n_tgt, const int64_t* tgt_index, const double* tgt_xyz,
                  int include_self, double* pot, double* field)
{
	#pragma omp parallel for schedule(dynamic, 4)
	for (int64_t t = 0; t < n_tgt; ++t) {
		const int64_t self = tgt_index ? tgt_index[t] : -1;
		const double* p = tgt_index ? S + 3 * self : tgt_xyz + 3 * t;
		double ph = 0, fx = 0, fy = 0, fz = 0;
		for (int64_t j = 0; j < n_src; ++j) {
			if (j == self && !include_self) continue;
			double dx = p[0] - S[3 * j], dy = p[1] - S[3 * j + 1], dz = p[2] - S[3 * j + 2];
			double r = sqrt(dx * dx + dy * dy + dz * dz);
			r = r > FORCE_EPS ? r : FORCE_EPS;
			ph += Q[j] / r;
			double r3 = r * (r * r);
			fx += Q[j] * dx / r3; fy += Q[j] * dy / r3; fz += Q[j] * dz / r3;
		}
		if (pot) pot[t] = ph;
		if (field) { field[3 * t] = fx; field[3 * t + 1] = fy; field[3 * t + 2] = fz; }
	}
}

/* Potential2d / Force2d, simple_math.h:119-128. */
void orc_direct2d(int64_t n_src, const double* S, const double* Q,
                  int64_t n_tgt, const int64_t* tgt_index, const double* tgt_xy,
                  int include_self, double* pot, double* field)
{
	#pragma omp parallel for schedule(dynamic, 4)
	for (int64_t t = 0; t < n_tgt; ++t) {
		const int64_t self = tgt_index ? tgt_index[t] : -1;
		const double* p = tgt_index ? S + 2 * self : tgt_xy + 2 * t;
		double ph = 0, fr = 0, fi = 0;
		for (int64_t j = 0; j < n_src; ++j) {
			if (j == self && !include_self) continue;
			double dx = p[0] - S[2 * j], dy = p[1] - S[2 * j + 1];
			double nz = dx * dx + dy * dy;
			nz = nz > FORCE_EPS2 ? nz : FORCE_EPS2;
			ph += Q[j] * 0.5 * log(nz);
			fr += Q[j] * dx / nz; fi += Q[j] * (-dy) / nz;
		}
		if (pot) pot[t] = ph;
		if (field) { field[2 * t] = fr; field[2 * t + 1] = fi; }
	}
}

/* ReadError, simple_math.h:342-402. */
void orc_read_error(int64_t n, int dim, const double* exact, const double* res, double* out)
{
	double mx = 0, e1 = 0, e2 = 0, l2 = 0;
	for (int64_t i = 0; i < n; ++i) {
		double de = 0, ax = 0;
		for (int a = 0; a < dim; ++a) {
			double d = exact[dim * i + a] - res[dim * i + a];
			de += d * d; ax += exact[dim * i + a] * exact[dim * i + a];
		}
		de = sqrt(de); ax = sqrt(ax);
		if (de > mx) mx = de;
		e1 += de; e2 += ax;
		l2 += (de / ax) * (de / ax);
	}
	out[0] = mx;
	out[1] = sqrt(l2 / (double)n);
	out[2] = e1 / e2;
}
