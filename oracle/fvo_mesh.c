/*
 * fvo_mesh.c — oracle restatement of the mesh layer (TEST INFRASTRUCTURE).
 *
 *   grd_read   spray-simulation/grd_read.f90:9-128
 *   grd_prepr  spray-simulation/grd_constr.f90:9-54 + gambit.vrt (element table)
 *   grd_conn   spray-simulation/grd_conn.f90:7-238
 *   grd_geom   spray-simulation/grd_geom.f90:9-225
 *   grd_math   spray-simulation/grd_math.f90:7-26
 *
 * pinned: golden dumps of the flang-built reference on cavity_3d.neu,
 * channel_2d.neu and generated boxes (tests/golden), closed forms on the
 * uniform hex (s=h^2, V=h^3, w_fp=1/2, l_pn=h), sum V = box volume, and
 * fvo_conn_hash == fvo_conn_literal on every fixture.
 */
#include "fvo.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define A1(type, n) ((type *)calloc((size_t)(n) + 1, sizeof(type)))

/* small = ten**(-20) as the flang build of precision.f90:8 evaluates it
   (bits 0x3BC79CA10C924224, one ulp above the nearest double to 1e-20);
   large = one/small (precision.f90:9). */
double fvo_small(void) { return 0x1.79ca10c924224p-67; }
double fvo_large(void) { return 1.0 / fvo_small(); }

/* ---- grd_math.f90 ---- */
static double vmag(const double *v, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + v[d] * v[d];
  return sqrt(s);
}
static double vdot(const double *a, const double *b, int n) {
  double s = 0.0;
  for (int d = 0; d < n; ++d) s = s + a[d] * b[d];
  return s;
}
/* cross_prod: c = v1(c1)*v2(c2) - v2(c1)*v1(c2), c1=(2,3,1), c2=(3,1,2) */
static void vcross(const double *a, const double *b, double *c) {
  static const int c1[3] = {1, 2, 0}, c2[3] = {2, 0, 1};
  double r[3];
  for (int d = 0; d < 3; ++d) r[d] = a[c1[d]] * b[c2[d]] - b[c1[d]] * a[c2[d]];
  c[0] = r[0]; c[1] = r[1]; c[2] = r[2];
}

/* ---- element table: gambit.vrt:4-47 (types 2..6; grd_prepr stops after
   type 6, grd_constr.f90:50, so the pyramid row is never loaded).  The wedge
   row gambit.vrt:30 lists local vertex 0; the third quad side is restated as
   3,1,4,6 (the only closing face consistent with rows 28-29). ---- */
typedef struct { int n_dx, n_vrt, n_fce; int fv[6][5]; } elem_t;
static const elem_t ELEM[7] = {
    {0, 0, 0, {{0}}},
    {0, 0, 0, {{0}}},
    {2, 4, 4, {{2, 1, 2}, {2, 2, 3}, {2, 3, 4}, {2, 4, 1}}},
    {2, 3, 3, {{2, 1, 2}, {2, 2, 3}, {2, 3, 1}}},
    {3, 8, 6, {{4, 1, 2, 6, 5}, {4, 2, 4, 8, 6}, {4, 4, 3, 7, 8}, {4, 3, 1, 5, 7}, {4, 2, 1, 3, 4}, {4, 5, 6, 8, 7}}},
    {3, 6, 5, {{4, 1, 2, 5, 4}, {4, 2, 3, 6, 5}, {4, 3, 1, 4, 6}, {3, 1, 3, 2}, {3, 4, 5, 6}}},
    {3, 4, 4, {{3, 2, 1, 3}, {3, 1, 2, 4}, {3, 2, 3, 4}, {3, 3, 1, 4}}},
};

int fvo_elem_nfce(int t) { return (t >= 2 && t <= 6) ? ELEM[t].n_fce : 0; }
int fvo_elem_face(int t, int ivf, int *out) {
  if (t < 2 || t > 6 || ivf < 1 || ivf > ELEM[t].n_fce) return 0;
  const int *f = ELEM[t].fv[ivf - 1];
  for (int k = 0; k < f[0]; ++k) out[k] = f[1 + k];
  return f[0];
}

fvo_mesh *fvo_mesh_raw(int n_dx, int n_vrt, const double *vx, int n_vol, const int *v_ty,
                       const int *v_vptr, const int *v_vrt, const int *bnd) {
  fvo_mesh *m = (fvo_mesh *)calloc(1, sizeof(fvo_mesh));
  m->n_dx = n_dx;
  m->n_vrt = n_vrt;
  m->n_vol = n_vol;
  m->vx = A1(double, 3 * (n_vrt + 1));
  for (int i = 1; i <= n_vrt; ++i)
    for (int d = 0; d < 3; ++d) m->vx[3 * i + d] = vx[3 * (i - 1) + d];
  m->v_ty = A1(int, n_vol);
  m->v_nvrt = A1(int, n_vol);
  m->v_nfce = A1(int, n_vol);
  m->v_vptr = A1(int, n_vol + 1);
  m->v_fptr = A1(int, n_vol + 1);
  int nv = 0, ns = 0;
  for (int i = 1; i <= n_vol; ++i) {
    m->v_ty[i] = v_ty[i - 1];
    m->v_nvrt[i] = v_vptr[i] - v_vptr[i - 1];
    m->v_nfce[i] = fvo_elem_nfce(m->v_ty[i]);
    nv += m->v_nvrt[i];
    ns += m->v_nfce[i];
  }
  m->n_slot = ns;
  m->v_vrt = A1(int, nv);
  m->v_bnd = A1(int, ns);
  m->v_fce = A1(int, ns);
  m->v_ng = A1(int, ns);
  m->v_vptr[1] = 1;
  m->v_fptr[1] = 1;
  for (int i = 1; i <= n_vol; ++i) {
    m->v_vptr[i + 1] = m->v_vptr[i] + m->v_nvrt[i];
    m->v_fptr[i + 1] = m->v_fptr[i] + m->v_nfce[i];
    for (int k = 0; k < m->v_nvrt[i]; ++k) m->v_vrt[m->v_vptr[i] + k] = v_vrt[v_vptr[i - 1] + k];
  }
  /* grd_read.f90:94-99: n_fce = (sum of cell faces + listed boundary faces)/2 */
  int nb = 0;
  for (int s = 1; s <= ns; ++s) {
    m->v_bnd[s] = bnd ? bnd[s - 1] : 0;
    if (m->v_bnd[s] > 0) ++nb;
  }
  m->n_fce = (ns + nb) / 2;
  return m;
}

void fvo_mesh_free(fvo_mesh *m) {
  if (!m) return;
  void *ptrs[] = {m->vx, m->v_ty, m->v_nvrt, m->v_nfce, m->v_vptr, m->v_vrt, m->v_fptr, m->v_bnd,
                  m->v_fce, m->v_ng, m->v_x, m->v_v, m->v_vr, m->v_sf, m->f_ip, m->f_in, m->f_ib,
                  m->f_bty, m->f_nvrt, m->f_vptr, m->f_vrt, m->f_x, m->f_n, m->f_dxp, m->f_dxn,
                  m->f_s, m->f_lpn, m->f_wfp};
  for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); ++k) free(ptrs[k]);
  free(m);
}

/* ---- grd_read.f90:115-128 bnd_ty ---- */
static int bnd_code(const char *name32) {
  char nm[64];
  int a = 0, n = 0;
  while (name32[a] == ' ') ++a;
  while (name32[a + n] && name32[a + n] != ' ' && name32[a + n] != '\n' && n < 63) {
    nm[n] = name32[a + n];
    ++n;
  }
  nm[n] = 0;
  const char *us = strchr(nm, '_');
  int bt = (us && us[1] >= '0' && us[1] <= '9') ? us[1] - '0' : 0;
  if (!strncmp(nm, "inlet", 5)) bt += 10;
  if (!strncmp(nm, "outlet", 6)) bt += 20;
  if (!strncmp(nm, "wall", 4)) bt += 30;
  if (!strncmp(nm, "symmetry", 8)) bt += 40;
  if (!strncmp(nm, "pressure", 8)) bt += 50;
  return bt;
}

/* first list-directed token of a line (blank/comma separated, '/' ends it) */
static void first_token(const char *line, char *tok) {
  int a = 0, n = 0;
  while (line[a] == ' ' || line[a] == '\t' || line[a] == ',') ++a;
  while (line[a + n] && !strchr(" \t,/\r\n", line[a + n]) && n < 31) {
    tok[n] = line[a + n];
    ++n;
  }
  tok[n] = 0;
}

static int next_ints(FILE *fp, char *line, int cap, int *out, int want, int *have, const char **cur) {
  /* read integers across continuation records */
  while (*have < want) {
    char *end;
    long v;
    while (**cur == ' ' || **cur == '\t') ++*cur;
    if (**cur == 0 || **cur == '\n' || **cur == '\r') {
      if (!fgets(line, cap, fp)) return -1;
      *cur = line;
      continue;
    }
    v = strtol(*cur, &end, 10);
    if (end == *cur) return -1;
    out[(*have)++] = (int)v;
    *cur = end;
  }
  return 0;
}

/* grd_read.f90:9-110 */
fvo_mesh *fvo_read_neu(const char *path, double l_sc) {
  FILE *fp = fopen(path, "r");
  if (!fp) return NULL;
  char line[512], tok[32];
  int n_vrt = 0, n_vol = 0, n_grp = 0, n_bs = 0, n_dx = 0, n_vc = 0, ibs = 0;
  double *vx = NULL;
  int *ty = NULL, *vptr = NULL, *vrt = NULL, *sptr = NULL, *bnd = NULL;
  int nvrt_tot = 0;
  while (fgets(line, sizeof line, fp)) {
    first_token(line, tok);
    if (!tok[0]) continue;
    if (!strcmp(tok, "NUMNP")) {
      if (!fgets(line, sizeof line, fp)) break;
      sscanf(line, "%d %d %d %d %d %d", &n_vrt, &n_vol, &n_grp, &n_bs, &n_dx, &n_vc);
      vx = (double *)calloc((size_t)3 * n_vrt, sizeof(double));
      ty = (int *)calloc((size_t)n_vol, sizeof(int));
      vptr = (int *)calloc((size_t)n_vol + 1, sizeof(int));
      vrt = (int *)calloc((size_t)8 * n_vol, sizeof(int));
    } else if (!strcmp(tok, "NODAL")) {
      for (int k = 0; k < n_vrt; ++k) {
        int idx;
        double x[3] = {0, 0, 0};
        if (!fgets(line, sizeof line, fp)) break;
        sscanf(line, "%d %lf %lf %lf", &idx, &x[0], &x[1], &x[2]);
        for (int d = 0; d < n_dx; ++d) vx[3 * (idx - 1) + d] = x[d] * l_sc;
      }
    } else if (!strcmp(tok, "ELEMENTS")) {
      /* records are in element order in every Gambit file; idx is honoured */
      int *tmp_n = (int *)calloc((size_t)n_vol + 1, sizeof(int));
      int *tmp_v = (int *)calloc((size_t)8 * (n_vol + 1), sizeof(int));
      for (int k = 0; k < n_vol; ++k) {
        int hdr[3], have = 0;
        const char *cur = "";
        if (next_ints(fp, line, sizeof line, hdr, 3, &have, &cur)) break;
        int idx = hdr[0];
        ty[idx - 1] = hdr[1];
        tmp_n[idx] = hdr[2];
        have = 0;
        if (next_ints(fp, line, sizeof line, &tmp_v[8 * idx], hdr[2], &have, &cur)) break;
      }
      vptr[0] = 0;
      for (int i = 1; i <= n_vol; ++i) {
        vptr[i] = vptr[i - 1] + tmp_n[i];
        for (int k = 0; k < tmp_n[i]; ++k) vrt[vptr[i - 1] + k] = tmp_v[8 * i + k];
      }
      nvrt_tot = vptr[n_vol];
      sptr = (int *)calloc((size_t)n_vol + 1, sizeof(int));
      for (int i = 1; i <= n_vol; ++i) sptr[i] = sptr[i - 1] + fvo_elem_nfce(ty[i - 1]);
      bnd = (int *)calloc((size_t)sptr[n_vol] + 1, sizeof(int));
      free(tmp_n);
      free(tmp_v);
    } else if (!strcmp(tok, "BOUNDARY")) {
      ibs++;
      if (!fgets(line, sizeof line, fp)) break;
      char name[33];
      memcpy(name, line, 32);
      name[32] = 0;
      int bty_file = 0, n = 0;
      sscanf(line + 32, "%d %d", &bty_file, &n);
      int code = bnd_code(name);
      for (int k = 0; k < n; ++k) {
        int ivol, et, ivf;
        if (!fgets(line, sizeof line, fp)) break;
        sscanf(line, "%d %d %d", &ivol, &et, &ivf);
        bnd[sptr[ivol - 1] + ivf - 1] = code;
      }
    }
    if (n_bs > 0 && ibs == n_bs) break;
  }
  fclose(fp);
  (void)nvrt_tot;
  (void)n_grp;
  (void)n_vc;
  if (!vx || !bnd) return NULL;
  fvo_mesh *m = fvo_mesh_raw(n_dx, n_vrt, vx, n_vol, ty, vptr, vrt, bnd);
  free(vx);
  free(ty);
  free(vptr);
  free(vrt);
  free(sptr);
  free(bnd);
  return m;
}

/* SplitMix64 (Steele, Lea, Flood 2014) — the jitter stream */
static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

fvo_mesh *fvo_gen_box(int nx, int ny, int nz, double lx, double ly, double lz, double jitter,
                      uint64_t seed, int lid_code, int wall_code) {
  int three = nz > 0;
  int nzv = three ? nz + 1 : 1;
  int n_vrt = (nx + 1) * (ny + 1) * nzv;
  int n_vol = nx * ny * (three ? nz : 1);
  double *vx = (double *)calloc((size_t)3 * n_vrt, sizeof(double));
  double h[3] = {lx / nx, ly / ny, three ? lz / nz : 0.0};
  uint64_t st = seed;
  for (int k = 0; k < nzv; ++k)
    for (int j = 0; j <= ny; ++j)
      for (int i = 0; i <= nx; ++i) {
        int v = i + (nx + 1) * (j + (ny + 1) * k);
        double *x = &vx[3 * v];
        x[0] = lx * (double)i / (double)nx;
        x[1] = ly * (double)j / (double)ny;
        x[2] = three ? lz * (double)k / (double)nz : 0.0;
        int interior = i > 0 && i < nx && j > 0 && j < ny && (!three || (k > 0 && k < nz));
        if (jitter > 0.0 && interior) {
          for (int d = 0; d < (three ? 3 : 2); ++d) {
            double u = (double)(splitmix64(&st) >> 11) * 0x1.0p-53;
            x[d] = x[d] + (2.0 * u - 1.0) * jitter * h[d];
          }
        }
      }
  int nv = three ? 8 : 4, nf = three ? 6 : 4;
  int *ty = (int *)malloc(sizeof(int) * (size_t)n_vol);
  int *vptr = (int *)malloc(sizeof(int) * ((size_t)n_vol + 1));
  int *vrt = (int *)malloc(sizeof(int) * (size_t)nv * n_vol);
  int *bnd = (int *)calloc((size_t)nf * n_vol, sizeof(int));
  vptr[0] = 0;
  int nzc = three ? nz : 1;
  for (int k = 0; k < nzc; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        int c = i + nx * (j + ny * k);
        ty[c] = three ? 4 : 2;
        vptr[c + 1] = vptr[c] + nv;
        int *w = &vrt[nv * c];
        int *b = &bnd[nf * c];
#define VID(a, bb, cc) (1 + (a) + (nx + 1) * ((bb) + (ny + 1) * (cc)))
        if (three) {
          w[0] = VID(i, j, k); w[1] = VID(i + 1, j, k); w[2] = VID(i, j + 1, k); w[3] = VID(i + 1, j + 1, k);
          w[4] = VID(i, j, k + 1); w[5] = VID(i + 1, j, k + 1); w[6] = VID(i, j + 1, k + 1); w[7] = VID(i + 1, j + 1, k + 1);
          if (j == 0) b[0] = wall_code;
          if (i == nx - 1) b[1] = wall_code;
          if (j == ny - 1) b[2] = wall_code;
          if (i == 0) b[3] = wall_code;
          if (k == 0) b[4] = wall_code;
          if (k == nz - 1) b[5] = lid_code;
        } else {
          w[0] = VID(i, j, 0); w[1] = VID(i + 1, j, 0); w[2] = VID(i + 1, j + 1, 0); w[3] = VID(i, j + 1, 0);
          if (j == 0) b[0] = wall_code;
          if (i == nx - 1) b[1] = wall_code;
          if (j == ny - 1) b[2] = lid_code;
          if (i == 0) b[3] = wall_code;
        }
#undef VID
      }
  fvo_mesh *m = fvo_mesh_raw(three ? 3 : 2, n_vrt, vx, n_vol, ty, vptr, vrt, bnd);
  free(vx); free(ty); free(vptr); free(vrt); free(bnd);
  return m;
}

/* ---- grd_conn ---- */
static void cell_face_vrt(const fvo_mesh *m, int ivol, int ivf, int *out, int *n) {
  int loc[5];
  *n = fvo_elem_face(m->v_ty[ivol], ivf, loc);
  for (int k = 0; k < *n; ++k) out[k] = m->v_vrt[m->v_vptr[ivol] + loc[k] - 1];
}

/* match_vrt, grd_conn.f90:161-181 */
static int match_vrt(const int *a, int na, const int *b, int nb) {
  if (na != nb) return 0;
  long sa = 0, sb = 0;
  for (int k = 0; k < na; ++k) { sa += a[k]; sb += b[k]; }
  if (sa != sb) return 0;
  int nm = 0;
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < na; ++j)
      if (a[i] == b[j]) { ++nm; break; }
  return nm == na;
}

static void alloc_faces(fvo_mesh *m) {
  int nf = m->n_fce;
  m->f_ip = A1(int, nf); m->f_in = A1(int, nf); m->f_ib = A1(int, nf); m->f_bty = A1(int, nf);
  m->f_nvrt = A1(int, nf); m->f_vptr = A1(int, nf + 1); m->f_vrt = A1(int, 4 * nf + 4);
}

/* grd_conn.f90:128-152 (fce_lst in ascending face order), c_vfce_lst
   :186-213 and c_neig_lst :218-238 */
static void cell_lists(fvo_mesh *m) {
  for (int s = 1; s <= m->n_slot; ++s) m->v_fce[s] = 0;
  for (int f = 1; f <= m->n_fce; ++f) {
    int ip = m->f_ip[f], in = m->f_in[f];
    for (int s = m->v_fptr[ip]; s < m->v_fptr[ip + 1]; ++s)
      if (m->v_fce[s] == 0) { m->v_fce[s] = f; break; }
    if (in > 0)
      for (int s = m->v_fptr[in]; s < m->v_fptr[in + 1]; ++s)
        if (m->v_fce[s] == 0) { m->v_fce[s] = f; break; }
  }
  for (int f = 1; f <= m->n_fce; ++f) {
    int ip = m->f_ip[f], in = m->f_in[f];
    for (int s = m->v_fptr[ip]; s < m->v_fptr[ip + 1]; ++s)
      if (m->v_fce[s] == 0) m->v_fce[s] = f;
    if (in > 0)
      for (int s = m->v_fptr[in]; s < m->v_fptr[in + 1]; ++s)
        if (m->v_fce[s] == 0) m->v_fce[s] = f;
  }
  for (int i = 1; i <= m->n_vol; ++i)
    for (int s = m->v_fptr[i]; s < m->v_fptr[i + 1]; ++s) {
      int f = m->v_fce[s];
      int ip = m->f_ip[f], in = m->f_in[f];
      m->v_ng[s] = 0;
      if (ip > 0 && ip != i) m->v_ng[s] = ip;
      else if (in > 0 && in != i) m->v_ng[s] = in;
    }
}

typedef struct { int ip, in, n, mtchd, copy; int v[4]; } flst_t;

static flst_t *build_flst(const fvo_mesh *m) {
  flst_t *L = (flst_t *)calloc((size_t)m->n_slot + 1, sizeof(flst_t));
  int k = 0;
  for (int i = 1; i <= m->n_vol; ++i)
    for (int ivf = 1; ivf <= m->v_nfce[i]; ++ivf) {
      ++k;
      L[k].ip = i;
      cell_face_vrt(m, i, ivf, L[k].v, &L[k].n);
    }
  return L;
}

static void emit_faces(fvo_mesh *m, flst_t *L) {
  alloc_faces(m);
  int f2 = 0, vp = 1;
  for (int k = 1; k <= m->n_slot; ++k) {
    if (L[k].copy) continue;
    ++f2;
    if (f2 > m->n_fce) break;
    m->f_nvrt[f2] = L[k].n;
    m->f_vptr[f2] = vp;
    for (int a = 0; a < L[k].n; ++a) m->f_vrt[vp + a] = L[k].v[a];
    vp += L[k].n;
    m->f_ip[f2] = L[k].ip;
    if (L[k].in > 0) m->f_in[f2] = L[k].in;
  }
  m->f_vptr[m->n_fce + 1] = vp;
}

/* grd_conn.f90:77-93, restated literally (O(n^2)) */
int fvo_conn_literal(fvo_mesh *m) {
  flst_t *L = build_flst(m);
  int n = m->n_slot;
  for (int a = 1; a <= n; ++a) {
    if (L[a].mtchd) continue;
    for (int b = 1; b <= n; ++b) {
      if (a == b || L[b].mtchd) continue;
      if (match_vrt(L[a].v, L[a].n, L[b].v, L[b].n)) {
        L[a].in = L[b].ip;
        L[a].mtchd = L[b].mtchd = 1;
        L[b].copy = 1;
        break;
      }
    }
  }
  emit_faces(m, L);
  free(L);
  cell_lists(m);
  return 0;
}

/* O(N) equivalent: hash on the sorted vertex tuple.  Scanning entries in
   (cell, local face) order, the first occurrence of a tuple is kept and its
   second occurrence becomes the copy, which is exactly the pairing the
   literal scan produces on a conforming mesh. */
typedef struct { int key[4]; int n; int idx; } hent_t;
static uint64_t hkey(const int *k, int n) {
  uint64_t h = 1469598103934665603ULL;
  for (int a = 0; a < n; ++a) { h ^= (uint64_t)(uint32_t)k[a]; h *= 1099511628211ULL; h ^= h >> 29; }
  return h;
}
int fvo_conn_hash(fvo_mesh *m) {
  flst_t *L = build_flst(m);
  int n = m->n_slot;
  size_t cap = 1;
  while (cap < (size_t)n * 2) cap <<= 1;
  hent_t *H = (hent_t *)calloc(cap, sizeof(hent_t));
  for (int a = 1; a <= n; ++a) {
    int key[4], kn = L[a].n;
    for (int t = 0; t < kn; ++t) key[t] = L[a].v[t];
    for (int x = 1; x < kn; ++x)
      for (int y = x; y > 0 && key[y - 1] > key[y]; --y) { int tt = key[y]; key[y] = key[y - 1]; key[y - 1] = tt; }
    size_t h = hkey(key, kn) & (cap - 1);
    for (;;) {
      if (H[h].idx == 0) {
        H[h].idx = a; H[h].n = kn;
        memcpy(H[h].key, key, sizeof key);
        break;
      }
      if (H[h].n == kn && !memcmp(H[h].key, key, sizeof(int) * kn)) {
        int f = H[h].idx;
        if (!L[f].mtchd) {
          L[f].in = L[a].ip;
          L[f].mtchd = L[a].mtchd = 1;
          L[a].copy = 1;
        }
        break;
      }
      h = (h + 1) & (cap - 1);
    }
  }
  free(H);
  emit_faces(m, L);
  free(L);
  cell_lists(m);
  return 0;
}

/* ---- grd_geom.f90:9-225 ---- */
int fvo_geom(fvo_mesh *m) {
  int nd = m->n_dx, nv = m->n_vol, nf = m->n_fce;
  double sm = fvo_small();
  m->v_x = A1(double, 3 * (nv + 1));
  m->v_v = A1(double, nv);
  m->v_vr = A1(double, nv);
  m->v_sf = A1(double, nv);
  m->f_x = A1(double, 3 * (nf + 1)); m->f_n = A1(double, 3 * (nf + 1));
  m->f_dxp = A1(double, 3 * (nf + 1)); m->f_dxn = A1(double, 3 * (nf + 1));
  m->f_s = A1(double, nf); m->f_lpn = A1(double, nf); m->f_wfp = A1(double, nf);
  m->n_bf = 0;
  for (int i = 1; i <= nv; ++i) {
    double acc[3] = {0, 0, 0};
    for (int k = m->v_vptr[i]; k < m->v_vptr[i + 1]; ++k)
      for (int d = 0; d < nd; ++d) acc[d] = acc[d] + m->vx[3 * m->v_vrt[k] + d];
    for (int d = 0; d < nd; ++d) m->v_x[3 * i + d] = acc[d] / (double)m->v_nvrt[i];
    /* boundary types and ib, :45-69 */
    for (int ivf = 1; ivf <= m->v_nfce[i]; ++ivf) {
      int code = m->v_bnd[m->v_fptr[i] + ivf - 1];
      if (code <= 0) continue;
      int lst[4], nl;
      cell_face_vrt(m, i, ivf, lst, &nl);
      for (int s = m->v_fptr[i]; s < m->v_fptr[i + 1]; ++s) {
        int f = m->v_fce[s];
        if (match_vrt(&m->f_vrt[m->f_vptr[f]], m->f_nvrt[f], lst, nl)) {
          m->f_bty[f] = code;
          m->n_bf++;
          m->f_ib[f] = nv + m->n_bf;
        }
      }
    }
  }
  m->n_vbf = nv + m->n_bf;
  for (int f = 1; f <= nf; ++f) {
    int ip = m->f_ip[f];
    double *xf = &m->f_x[3 * f], *n = &m->f_n[3 * f];
    const int *fv = &m->f_vrt[m->f_vptr[f]];
    if (nd == 2) {
      const double *x1 = &m->vx[3 * fv[0]], *x2 = &m->vx[3 * fv[1]];
      double e[3] = {x2[0] - x1[0], x2[1] - x1[1], 0.0};
      double s0 = vmag(e, 3);
      double z[3] = {0.0, 0.0, 1.0}, c[3];
      vcross(e, z, c);
      double mg = vmag(c, 2);
      n[0] = c[0] / mg;
      n[1] = c[1] / mg;
      m->f_s[f] = s0;
      xf[0] = 0.5 * (x1[0] + x2[0]);
      xf[1] = 0.5 * (x1[1] + x2[1]);
    } else {
      int nfv = m->f_nvrt[f];
      double vec[5][3] = {{0}}, s0 = 0.0;
      const double *x1 = &m->vx[3 * fv[0]];
      for (int k = 2; k <= nfv; ++k) {
        const double *x2 = &m->vx[3 * fv[k - 1]];
        for (int d = 0; d < 3; ++d) vec[k][d] = x2[d] - x1[d];
        if (k > 2) {
          double cr[3];
          vcross(vec[k - 1], vec[k], cr);
          for (int d = 0; d < 3; ++d) vec[1][d] = 0.5 * cr[d];
          double s1 = vmag(vec[1], 3);
          for (int d = 0; d < 3; ++d) vec[0][d] = vec[0][d] + vec[1][d] / s1;
          s0 = s0 + s1;
        }
      }
      for (int d = 0; d < 3; ++d) n[d] = vec[0][d] / (double)(nfv - 2);
      m->f_s[f] = s0;
      double acc[3] = {0, 0, 0};
      for (int k = 0; k < nfv; ++k)
        for (int d = 0; d < 3; ++d) acc[d] = acc[d] + m->vx[3 * fv[k] + d];
      for (int d = 0; d < 3; ++d) xf[d] = acc[d] / (double)nfv;
    }
    /* volume by x-flux, :152-153 */
    double vp = xf[0] * m->f_s[f] * n[0];
    m->v_v[ip] = m->v_v[ip] + vp;
    const double *xP = &m->v_x[3 * ip];
    double t[3], xpa[3];
    for (int d = 0; d < nd; ++d) t[d] = xf[d] - xP[d];
    double dd = vdot(t, n, nd);
    for (int d = 0; d < nd; ++d) xpa[d] = xf[d] - n[d] * dd;
    for (int d = 0; d < nd; ++d) m->f_dxp[3 * f + d] = xpa[d] - xP[d];
    int in = m->f_in[f];
    if (in > 0) {
      double vpn = xf[0] * m->f_s[f] * n[0];
      m->v_v[in] = m->v_v[in] - vpn;
      const double *xN = &m->v_x[3 * in];
      double xna[3], dpn[3], dfp[3];
      for (int d = 0; d < nd; ++d) t[d] = xf[d] - xN[d];
      double dn = vdot(t, n, nd);
      for (int d = 0; d < nd; ++d) xna[d] = xf[d] - n[d] * dn;
      for (int d = 0; d < nd; ++d) m->f_dxn[3 * f + d] = xna[d] - xN[d];
      for (int d = 0; d < nd; ++d) dpn[d] = xna[d] - xpa[d];
      m->f_lpn[f] = vmag(dpn, nd);
      for (int d = 0; d < nd; ++d) dfp[d] = xf[d] - xpa[d];
      m->f_wfp[f] = vmag(dfp, nd) / m->f_lpn[f];
    } else {
      double dfp[3];
      for (int d = 0; d < nd; ++d) dfp[d] = xf[d] - xpa[d];
      m->f_lpn[f] = vmag(dfp, nd);
    }
  }
  for (int i = 1; i <= nv; ++i) m->v_vr[i] = 1.0 / m->v_v[i];
  (void)sm;
  return 0;
}
