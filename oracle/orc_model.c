/*
 * orc_model.c -- input parsing, lattice/general model construction and the
 * model-phase RNG draws, restated from the reference.  Test infrastructure
 * only (see lsqt_oracle.h).
 *
 *   Model::Model                      src/model.cu:30-65
 *   initialize_parameters / verify    src/model.cu:243-310 / 128-241
 *   energy.in / local_orbitals / time src/model.cu:312-334 / 336-361 / 363-384
 *   initialize_state                  src/model.cu:77-109
 *   lattice model                     src/model_lattice.cu:148-402
 *   vacancies                         src/model_lattice.cu:26-146
 *   general model                     src/model_general.cu:24-185
 *   Anderson disorder                 src/anderson.cu:22-29
 *   charged impurities                src/charge.cu:25-231
 *
 * Deliberate difference: where the reference would loop forever or index
 * out of bounds on an invalid input (a charged-impurity cell count of 0,
 * more hoppings than max_neighbor, an odd site count with spin, M < 2), the
 * oracle reports an error instead.
 */
#define _DEFAULT_SOURCE
#include "lsqt_oracle.h"

#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.141592653589793

/* ---- whitespace token stream with C++ `>>` semantics for valid input ---- */
typedef struct {
  char* buf;
  size_t len, pos;
  int fail;
} tokstream;

static int ts_open(tokstream* ts, const char* path)
{
  memset(ts, 0, sizeof(*ts));
  FILE* f = fopen(path, "rb");
  if (!f)
    return 0;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  ts->buf = (char*)malloc((size_t)n + 1);
  ts->len = fread(ts->buf, 1, (size_t)n, f);
  ts->buf[ts->len] = 0;
  fclose(f);
  return 1;
}

static void ts_close(tokstream* ts)
{
  free(ts->buf);
  ts->buf = NULL;
}

static int is_space(char c)
{
  return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f';
}

/* next token as [start, end) */
static int ts_token(tokstream* ts, size_t* start, size_t* end)
{
  if (ts->fail)
    return 0;
  while (ts->pos < ts->len && is_space(ts->buf[ts->pos]))
    ++ts->pos;
  if (ts->pos >= ts->len) {
    ts->fail = 1;
    return 0;
  }
  *start = ts->pos;
  while (ts->pos < ts->len && !is_space(ts->buf[ts->pos]))
    ++ts->pos;
  *end = ts->pos;
  return 1;
}

static int ts_double(tokstream* ts, double* out)
{
  size_t s, e;
  if (!ts_token(ts, &s, &e)) {
    *out = 0.0;
    return 0;
  }
  char save = ts->buf[e];
  ts->buf[e] = 0;
  char* endp;
  double v = strtod(ts->buf + s, &endp);
  ts->buf[e] = save;
  if (endp == ts->buf + s) {
    ts->fail = 1;
    *out = 0.0;
    return 0;
  }
  *out = v;
  return 1;
}

static int ts_int(tokstream* ts, int* out)
{
  size_t s, e;
  if (!ts_token(ts, &s, &e)) {
    *out = 0;
    return 0;
  }
  char save = ts->buf[e];
  ts->buf[e] = 0;
  char* endp;
  long v = strtol(ts->buf + s, &endp, 10);
  ts->buf[e] = save;
  if (endp == ts->buf + s) {
    ts->fail = 1;
    *out = 0;
    return 0;
  }
  *out = (int)v;
  return 1;
}

static int ts_word(tokstream* ts, char* out, size_t cap)
{
  size_t s, e;
  out[0] = 0;
  if (!ts_token(ts, &s, &e))
    return 0;
  size_t n = e - s < cap - 1 ? e - s : cap - 1;
  memcpy(out, ts->buf + s, n);
  out[n] = 0;
  return 1;
}

static int fail(orc_model* m, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(m->error, sizeof(m->error), fmt, ap);
  va_end(ap);
  return 1;
}

static void path_join(char* out, size_t cap, const char* dir, const char* name)
{
  snprintf(out, cap, "%s/%s", dir, name);
}

/* ---- para.in (src/model.cu:243-310) and its checks (:128-241) ----------- */
static int parse_parameters(orc_model* m)
{
  char path[4200];
  path_join(path, sizeof(path), m->input_dir, "para.in");
  FILE* f = fopen(path, "r");
  if (!f)
    return fail(m, "Error: cannot open %s", path);
  char line[4096];
  while (fgets(line, sizeof(line), f)) {
    size_t L = strlen(line);
    while (L && (line[L - 1] == '\n'))
      line[--L] = 0;
    tokstream ts = {line, L, 0, 0};
    char tok[256];
    if (!ts_word(&ts, tok, sizeof(tok)))
      continue; /* blank line */
    int iv;
    if (!strcmp(tok, "model")) {
      ts_int(&ts, &iv);
      m->use_lattice_model = iv != 0;
    } else if (!strcmp(tok, "anderson_disorder")) {
      m->has_anderson = 1;
      ts_double(&ts, &m->anderson_W);
    } else if (!strcmp(tok, "charged_impurity")) {
      m->has_charge = 1;
      ts_int(&ts, &m->charge_Ni);
      ts_double(&ts, &m->charge_W);
      ts_double(&ts, &m->charge_xi);
    } else if (!strcmp(tok, "vacancy_disorder")) {
      m->has_vacancy = 1;
      ts_int(&ts, &m->number_of_vacancies);
    } else if (!strcmp(tok, "calculate_vac0")) {
      m->calculate_vac0 = 1;
    } else if (!strcmp(tok, "calculate_vac")) {
      m->calculate_vac = 1;
    } else if (!strcmp(tok, "calculate_msd")) {
      m->calculate_msd = 1;
    } else if (!strcmp(tok, "calculate_spin")) {
      m->calculate_spin = 1;
    } else if (!strcmp(tok, "calculate_ldos")) {
      m->calculate_ldos = 1;
    } else if (!strcmp(tok, "number_of_random_vectors")) {
      ts_int(&ts, &m->number_of_random_vectors);
    } else if (!strcmp(tok, "number_of_moments")) {
      ts_int(&ts, &m->number_of_moments);
    } else if (!strcmp(tok, "energy_max")) {
      ts_double(&ts, &m->energy_max);
    } else {
      fclose(f);
      return fail(m, "Error: Unknown identifier in %s/para.in: %s", m->input_dir, line);
    }
  }
  fclose(f);

  m->requires_time = m->calculate_vac || m->calculate_msd || m->calculate_spin;
  if (m->use_lattice_model) {
    if (m->calculate_spin)
      return fail(m, "Error: lattice model does not support spin calculations yet");
    if (m->has_vacancy && m->number_of_vacancies <= 0)
      return fail(m, "Error: number of vacancies should > 0");
  } else {
    if (m->has_anderson)
      return fail(m, "Error: General model does not allowed to add Anderson disorder");
    if (m->has_vacancy)
      return fail(m, "Error: General model does not allowed to add vacancy disorder");
    if (m->has_charge)
      return fail(m, "Error: General model does not allowed to add charged impurities");
  }
  if (m->calculate_spin && m->calculate_vac0)
    return fail(m, "Error: spin and VAC0 cannot be calculated together");
  if (m->calculate_spin && m->calculate_vac)
    return fail(m, "Error: spin and VAC cannot be calculated together");
  if (m->calculate_spin && m->calculate_msd)
    return fail(m, "Error: spin and MSD cannot be calculated together");
  if (m->number_of_random_vectors <= 0)
    return fail(m, "Error: Number of random vectors should > 0");
  if (m->number_of_moments <= 0)
    return fail(m, "Error: Number of moments should > 0");
  if (m->energy_max <= 0)
    return fail(m, "Error: Energy maximum should > 0");
  if (m->number_of_moments < 2)
    return fail(m, "Error: Number of moments should be >= 2");
  return 0;
}

/* ---- lattice model ------------------------------------------------------ */
/* src/model_lattice.cu:148-163: one wrap per direction, x slowest */
static int find_index(int nx, int ny, int nz, int Nx, int Ny, int Nz, int m, int N_orbital)
{
  if (nx < 0)
    nx += Nx;
  if (nx >= Nx)
    nx -= Nx;
  if (ny < 0)
    ny += Ny;
  if (ny >= Ny)
    ny -= Ny;
  if (nz < 0)
    nz += Nz;
  if (nz >= Nz)
    nz -= Nz;
  return ((nx * Ny + ny) * Nz + nz) * N_orbital + m;
}

/* src/model_lattice.cu:26-43: a full shuffle of [0, max_value) with one
 * draw per position, of which the first total_number entries are kept */
static void random_permutation_prefix(orc_mt19937* g, int max_value, int total_number,
                                      int* out)
{
  int* p = (int*)malloc(sizeof(int) * (size_t)(max_value > 0 ? max_value : 1));
  for (int i = 0; i < max_value; ++i)
    p[i] = i;
  for (int i = 0; i < max_value; ++i) {
    int j = orc_uniform_int_0_intmax(g) % (max_value - i) + i;
    int t = p[i];
    p[i] = p[j];
    p[j] = t;
  }
  for (int i = 0; i < total_number; ++i)
    out[i] = p[i];
  free(p);
}

/* src/model_lattice.cu:70-146 */
static int add_vacancies(orc_model* m)
{
  int N0 = m->number_of_atoms, K = m->max_neighbor;
  if (m->number_of_vacancies > N0)
    return fail(m, "Error: more vacancies than sites");
  int* is_vac = (int*)calloc((size_t)N0, sizeof(int));
  int* idx = (int*)malloc(sizeof(int) * (size_t)m->number_of_vacancies);
  random_permutation_prefix(&m->generator, N0, m->number_of_vacancies, idx);
  for (int n = 0; n < m->number_of_vacancies; ++n)
    is_vac[idx[n]] = 1;
  free(idx);
  int* new_index = (int*)malloc(sizeof(int) * (size_t)N0);
  int count = 0;
  for (int n = 0; n < N0; ++n)
    if (!is_vac[n])
      new_index[n] = count++;

  int N1 = N0 - m->number_of_vacancies;
  size_t P1 = (size_t)N1 * (size_t)K;
  int* nn = (int*)calloc((size_t)(N1 > 0 ? N1 : 1), sizeof(int));
  int* nl = (int*)calloc(P1 ? P1 : 1, sizeof(int));
  double* hr = (double*)calloc(P1 ? P1 : 1, sizeof(double));
  double* hi = (double*)calloc(P1 ? P1 : 1, sizeof(double));
  double* xx = (double*)calloc(P1 ? P1 : 1, sizeof(double));
  int ca = 0;
  for (int n = 0; n < N0; ++n) {
    if (is_vac[n])
      continue;
    int cn = 0;
    for (int j = 0; j < m->neighbor_number[n]; ++j) {
      size_t old = (size_t)n * K + j;
      int k = m->neighbor_list[old];
      if (!is_vac[k]) {
        size_t nw = (size_t)ca * K + cn;
        nl[nw] = new_index[k];
        hr[nw] = m->hopping_real[old];
        hi[nw] = m->hopping_imag[old];
        xx[nw] = m->xx[old];
        ++cn;
      }
    }
    nn[ca] = cn;
    ++ca;
  }
  free(m->neighbor_number);
  free(m->neighbor_list);
  free(m->hopping_real);
  free(m->hopping_imag);
  free(m->xx);
  m->neighbor_number = nn;
  m->neighbor_list = nl;
  m->hopping_real = hr;
  m->hopping_imag = hi;
  m->xx = xx;
  m->number_of_atoms = N1;
  m->number_of_pairs = (int)P1;
  free(is_vac);
  free(new_index);
  return 0;
}

/* ---- charged impurities (src/charge.cu) --------------------------------- */
typedef struct {
  int Nx, Ny, Nz, Nxyz;
  double rc, rc2;
  int *cell_count, *cell_count_sum, *cell_contents;
} cells;

static int cell_id(const cells* c, double x, double y, double z, int* pnx, int* pny, int* pnz)
{
  int nx = (int)floor(x / c->rc);
  int ny = (int)floor(y / c->rc);
  int nz = (int)floor(z / c->rc);
  while (nx < 0)
    nx += c->Nx;
  while (nx >= c->Nx)
    nx -= c->Nx;
  while (ny < 0)
    ny += c->Ny;
  while (ny >= c->Ny)
    ny -= c->Ny;
  while (nz < 0)
    nz += c->Nz;
  while (nz >= c->Nz)
    nz -= c->Nz;
  if (pnx) {
    *pnx = nx;
    *pny = ny;
    *pnz = nz;
  }
  return nx + c->Nx * ny + c->Nx * c->Ny * nz;
}

static int neighbor_cell(const cells* c, int nx, int ny, int nz, int nxyz, int i, int j, int k)
{
  int nb = nxyz + k * c->Nx * c->Ny + j * c->Nx + i;
  if (nx + i < 0)
    nb += c->Nx;
  if (nx + i >= c->Nx)
    nb -= c->Nx;
  if (ny + j < 0)
    nb += c->Ny * c->Nx;
  if (ny + j >= c->Ny)
    nb -= c->Ny * c->Nx;
  if (nz + k < 0)
    nb += c->Nxyz;
  if (nz + k >= c->Nz)
    nb -= c->Nxyz;
  return nb;
}

static int add_charges(orc_model* m, const int pbc[3], const double box[3], const double* x,
                       const double* y, const double* z)
{
  int N = m->number_of_atoms, Ni = m->charge_Ni;
  cells c;
  memset(&c, 0, sizeof(c));
  c.rc = 5.0 * m->charge_xi;
  c.rc2 = c.rc * c.rc;
  if (Ni < 0 || Ni > N)
    return fail(m, "Error: invalid number of charged impurities");
  int* imp = (int*)malloc(sizeof(int) * (size_t)(Ni > 0 ? Ni : 1));
  double* str = (double*)malloc(sizeof(double) * (size_t)(Ni > 0 ? Ni : 1));
  random_permutation_prefix(&m->generator, N, Ni, imp);
  double W2 = m->charge_W * 0.5;
  for (int i = 0; i < Ni; ++i)
    str[i] = orc_uniform_real(&m->generator, -W2, W2);

  c.Nx = pbc[0] ? (int)floor(box[0] / c.rc) : 1;
  c.Ny = pbc[1] ? (int)floor(box[1] / c.rc) : 1;
  c.Nz = pbc[2] ? (int)floor(box[2] / c.rc) : 1;
  c.Nxyz = c.Nx * c.Ny * c.Nz;
  if (c.Nx <= 0 || c.Ny <= 0 || c.Nz <= 0) {
    free(imp);
    free(str);
    return fail(m, "Error: charged impurity range 5*xi exceeds the box length");
  }
  c.cell_count = (int*)calloc((size_t)c.Nxyz, sizeof(int));
  c.cell_count_sum = (int*)calloc((size_t)c.Nxyz, sizeof(int));
  c.cell_contents = (int*)calloc((size_t)(N > 0 ? N : 1), sizeof(int));
  for (int i = 0; i < N; ++i)
    ++c.cell_count[cell_id(&c, x[i], y[i], z[i], NULL, NULL, NULL)];
  for (int i = 1; i < c.Nxyz; ++i)
    c.cell_count_sum[i] = c.cell_count[i - 1] + c.cell_count_sum[i - 1];
  memset(c.cell_count, 0, sizeof(int) * (size_t)c.Nxyz);
  for (int i = 0; i < N; ++i) {
    int id = cell_id(&c, x[i], y[i], z[i], NULL, NULL, NULL);
    c.cell_contents[c.cell_count_sum[id] + c.cell_count[id]] = i;
    ++c.cell_count[id];
  }

  double xi_factor = -0.5 / (m->charge_xi * m->charge_xi);
  double half[3];
  for (int d = 0; d < 3; ++d)
    half[d] = box[d] * 0.5;
  for (int n = 0; n < N; ++n)
    m->potential[n] = 0.0;
  for (int ni = 0; ni < Ni; ++ni) {
    int n1 = imp[ni];
    double x1 = x[n1], y1 = y[n1], z1 = z[n1];
    int nx, ny, nz;
    int nxyz = cell_id(&c, x1, y1, z1, &nx, &ny, &nz);
    int klim = pbc[2] ? 1 : 0, jlim = pbc[1] ? 1 : 0, ilim = pbc[0] ? 1 : 0;
    for (int k = -klim; k < klim + 1; ++k)
      for (int j = -jlim; j < jlim + 1; ++j)
        for (int i = -ilim; i < ilim + 1; ++i) {
          int nb = neighbor_cell(&c, nx, ny, nz, nxyz, i, j, k);
          for (int q = 0; q < c.cell_count[nb]; ++q) {
            int n2 = c.cell_contents[c.cell_count_sum[nb] + q];
            double r[3] = {x[n2] - x1, y[n2] - y1, z[n2] - z1};
            double d2 = 0.0;
            for (int d = 0; d < 3; ++d) {
              r[d] = fabs(r[d]);
              if (pbc[d] == 1 && r[d] > half[d])
                r[d] = box[d] - r[d];
              d2 += r[d] * r[d];
            }
            if (d2 > c.rc2)
              continue;
            m->potential[n2] += str[ni] * exp(d2 * xi_factor);
          }
        }
  }
  free(imp);
  free(str);
  free(c.cell_count);
  free(c.cell_count_sum);
  free(c.cell_contents);
  return 0;
}

/* src/model_lattice.cu:165-402 */
static int build_lattice(orc_model* m)
{
  char path[4200];
  path_join(path, sizeof(path), m->input_dir, "lattice.in");
  tokstream ts;
  if (!ts_open(&ts, path))
    return fail(m, "Could not open %s", path);
  int Nc[3], pbc[3], td, N_orb, K;
  double lc[3], box[3];
  for (int d = 0; d < 3; ++d)
    ts_int(&ts, &Nc[d]);
  const char* axis[3] = {"x", "y", "z"};
  for (int d = 0; d < 3; ++d)
    if (Nc[d] <= 0) {
      ts_close(&ts);
      return fail(m, "Error: Number of cells in the %s direction should > 0", axis[d]);
    }
  for (int d = 0; d < 3; ++d)
    ts_int(&ts, &pbc[d]);
  ts_int(&ts, &td);
  for (int d = 0; d < 3; ++d)
    if (pbc[d] != 0 && pbc[d] != 1) {
      ts_close(&ts);
      return fail(m, "Error: %s direction has wrong boundary", axis[d]);
    }
  if (td < 0 || td > 2) {
    ts_close(&ts);
    return fail(m, "Error: wrong transport direction");
  }
  if (pbc[td] != 1) {
    ts_close(&ts);
    return fail(m, "Error: transport direction must be periodic");
  }
  for (int d = 0; d < 3; ++d)
    ts_double(&ts, &lc[d]);
  for (int d = 0; d < 3; ++d)
    if (lc[d] <= 0) {
      ts_close(&ts);
      return fail(m, "Error: lattice constant in %s direction < 0", axis[d]);
    }
  for (int d = 0; d < 3; ++d)
    box[d] = lc[d] * Nc[d];
  m->volume = box[0] * box[1] * box[2];
  ts_int(&ts, &N_orb);
  ts_int(&ts, &K);
  if (N_orb <= 0) {
    ts_close(&ts);
    return fail(m, "Error: number of orbitals per cell should > 0");
  }
  if (K <= 0) {
    ts_close(&ts);
    return fail(m, "Error: maximum number of hoppings per orbital should > 0");
  }
  long long Nll = (long long)N_orb * Nc[0] * Nc[1] * Nc[2];
  if (Nll * K > INT_MAX) {
    ts_close(&ts);
    return fail(m, "Error: system too large for int32 indexing");
  }
  int N = (int)Nll;
  m->number_of_atoms = N;
  m->max_neighbor = K;
  m->number_of_pairs = N * K;
  size_t P = (size_t)N * K;
  m->neighbor_number = (int*)calloc((size_t)N, sizeof(int));
  m->neighbor_list = (int*)calloc(P, sizeof(int));
  m->hopping_real = (double*)calloc(P, sizeof(double));
  m->hopping_imag = (double*)calloc(P, sizeof(double));
  m->xx = (double*)calloc(P, sizeof(double));
  m->potential = (double*)calloc((size_t)N, sizeof(double));
  double *x = NULL, *y = NULL, *z = NULL;
  if (m->has_charge) {
    x = (double*)calloc((size_t)N, sizeof(double));
    y = (double*)calloc((size_t)N, sizeof(double));
    z = (double*)calloc((size_t)N, sizeof(double));
  }
  double* xc = (double*)calloc((size_t)N_orb, sizeof(double));
  double* yc = (double*)calloc((size_t)N_orb, sizeof(double));
  double* zc = (double*)calloc((size_t)N_orb, sizeof(double));
  for (int n = 0; n < N_orb; ++n) {
    ts_double(&ts, &xc[n]);
    ts_double(&ts, &yc[n]);
    ts_double(&ts, &zc[n]);
  }
  int nh_cell = N_orb * K;
  int* hidx[4];
  for (int q = 0; q < 4; ++q)
    hidx[q] = (int*)calloc((size_t)nh_cell, sizeof(int));
  double* hdat[2];
  for (int q = 0; q < 2; ++q)
    hdat[q] = (double*)calloc((size_t)nh_cell, sizeof(double));
  int* nhop = (int*)calloc((size_t)N_orb, sizeof(int));
  int rc = 0;
  for (int o = 0; o < N_orb && !rc; ++o) {
    ts_int(&ts, &nhop[o]);
    if (nhop[o] > K || nhop[o] < 0) {
      rc = fail(m, "Error: orbital %d has more hoppings than the maximum", o);
      break;
    }
    for (int h = 0; h < nhop[o]; ++h) {
      int q = o * K + h;
      ts_int(&ts, &hidx[0][q]);
      ts_int(&ts, &hidx[1][q]);
      ts_int(&ts, &hidx[2][q]);
      ts_int(&ts, &hidx[3][q]);
      ts_double(&ts, &hdat[0][q]);
      ts_double(&ts, &hdat[1][q]);
    }
  }
  ts_close(&ts);

  for (int nx1 = 0; nx1 < Nc[0] && !rc; ++nx1)
    for (int ny1 = 0; ny1 < Nc[1]; ++ny1)
      for (int nz1 = 0; nz1 < Nc[2]; ++nz1)
        for (int o = 0; o < N_orb; ++o) {
          int n1 = find_index(nx1, ny1, nz1, Nc[0], Nc[1], Nc[2], o, N_orb);
          if (m->has_charge) {
            x[n1] = xc[o] + lc[0] * nx1;
            y[n1] = yc[o] + lc[1] * ny1;
            z[n1] = zc[o] + lc[2] * nz1;
          }
          int count = 0;
          for (int i = 0; i < nhop[o]; ++i) {
            size_t slot = (size_t)n1 * K + count;
            int q = o * K + i;
            int nx2 = hidx[0][q] + nx1, ny2 = hidx[1][q] + ny1, nz2 = hidx[2][q] + nz1;
            int skip_x = !pbc[0] && (nx2 < 0 || nx2 >= Nc[0]);
            int skip_y = !pbc[1] && (ny2 < 0 || ny2 >= Nc[1]);
            int skip_z = !pbc[2] && (nz2 < 0 || nz2 >= Nc[2]);
            if (skip_x || skip_y || skip_z)
              continue;
            m->neighbor_list[slot] =
              find_index(nx2, ny2, nz2, Nc[0], Nc[1], Nc[2], hidx[3][q], N_orb);
            /* the orbital offset is taken from the x coordinates whatever the
             * transport direction (src/model_lattice.cu:373-375) */
            double x12 = lc[td] * hidx[td][q];
            x12 += xc[hidx[3][q]] - xc[o];
            m->xx[slot] = x12;
            m->hopping_real[slot] = hdat[0][q];
            m->hopping_imag[slot] = hdat[1][q];
            ++count;
          }
          m->neighbor_number[n1] = count;
        }
  for (int q = 0; q < 4; ++q)
    free(hidx[q]);
  for (int q = 0; q < 2; ++q)
    free(hdat[q]);
  free(nhop);
  free(xc);
  free(yc);
  free(zc);

  if (!rc && m->has_vacancy)
    rc = add_vacancies(m);
  if (!rc && m->has_anderson) {
    double W2 = m->anderson_W * 0.5;
    for (int n = 0; n < m->number_of_atoms; ++n)
      m->potential[n] = orc_uniform_real(&m->generator, -W2, W2);
  }
  /* positions stay indexed by the pristine lattice even after vacancies
   * (src/model_lattice.cu:389-399) */
  if (!rc && m->has_charge)
    rc = add_charges(m, pbc, box, x, y, z);
  free(x);
  free(y);
  free(z);
  return rc;
}

/* ---- general model (src/model_general.cu) ------------------------------- */
static double reduce_distance(double d, double box)
{
  if (d > box / 2.0)
    return d - box;
  if (d < -box / 2.0)
    return d + box;
  return d;
}

static int build_general(orc_model* m)
{
  char path[4200];
  tokstream ts;
  path_join(path, sizeof(path), m->input_dir, "neighbor.in");
  if (!ts_open(&ts, path))
    return fail(m, "Error: cannot open %s", path);
  int N, K;
  ts_int(&ts, &N);
  ts_int(&ts, &K);
  if (N <= 0 || K <= 0 || (long long)N * K > INT_MAX) {
    ts_close(&ts);
    return fail(m, "Error: invalid neighbor.in header");
  }
  m->number_of_atoms = N;
  m->max_neighbor = K;
  m->number_of_pairs = N * K;
  size_t P = (size_t)N * K;
  m->neighbor_number = (int*)calloc((size_t)N, sizeof(int));
  m->neighbor_list = (int*)calloc(P, sizeof(int));
  for (int n = 0; n < N; ++n) {
    ts_int(&ts, &m->neighbor_number[n]);
    if (m->neighbor_number[n] > K || m->neighbor_number[n] < 0) {
      ts_close(&ts);
      return fail(m, "Error: site %d has more neighbors than the maximum", n);
    }
    for (int j = 0; j < m->neighbor_number[n]; ++j) {
      int* e = &m->neighbor_list[(size_t)n * K + j];
      ts_int(&ts, e);
      if (*e < 0 || *e >= N) {
        ts_close(&ts);
        return fail(m, "Error: neighbor index out of range at site %d", n);
      }
    }
  }
  ts_close(&ts);

  path_join(path, sizeof(path), m->input_dir, "position.in");
  if (!ts_open(&ts, path))
    return fail(m, "Error: cannot open %s", path);
  double box;
  ts_double(&ts, &box);
  ts_double(&ts, &m->volume);
  double* x = (double*)calloc((size_t)N, sizeof(double));
  for (int i = 0; i < N; ++i)
    ts_double(&ts, &x[i]);
  ts_close(&ts);
  m->xx = (double*)calloc(P, sizeof(double));
  for (int n = 0; n < N; ++n)
    for (int j = 0; j < m->neighbor_number[n]; ++j) {
      size_t q = (size_t)n * K + j;
      m->xx[q] = reduce_distance(x[m->neighbor_list[q]] - x[n], box);
    }
  free(x);

  m->potential = (double*)calloc((size_t)N, sizeof(double));
  path_join(path, sizeof(path), m->input_dir, "potential.in");
  if (ts_open(&ts, path)) {
    for (int n = 0; n < N; ++n)
      ts_double(&ts, &m->potential[n]);
    ts_close(&ts);
  }

  m->hopping_real = (double*)calloc(P, sizeof(double));
  m->hopping_imag = (double*)calloc(P, sizeof(double));
  int type = 3;
  path_join(path, sizeof(path), m->input_dir, "hopping.in");
  if (ts_open(&ts, path)) {
    char word[64];
    ts_word(&ts, word, sizeof(word));
    if (!strcmp(word, "complex"))
      type = 1;
    else if (!strcmp(word, "real"))
      type = 2;
    else {
      ts_close(&ts);
      return fail(m, "- Hoppings can only be real or complex");
    }
  }
  for (int n = 0; n < N; ++n)
    for (int j = 0; j < m->neighbor_number[n]; ++j) {
      size_t q = (size_t)n * K + j;
      if (type < 3)
        ts_double(&ts, &m->hopping_real[q]);
      else
        m->hopping_real[q] = -1.0;
      if (type == 1)
        ts_double(&ts, &m->hopping_imag[q]);
      else
        m->hopping_imag[q] = 0.0;
    }
  if (type < 3)
    ts_close(&ts);
  return 0;
}

/* ---- count-then-values files (src/model.cu:312-384) --------------------- */
static int read_counted(orc_model* m, const char* name, int* count, double** values)
{
  char path[4200];
  path_join(path, sizeof(path), m->input_dir, name);
  tokstream ts;
  if (!ts_open(&ts, path))
    return fail(m, "Error: cannot open %s", path);
  ts_int(&ts, count);
  if (*count < 0)
    *count = 0;
  *values = (double*)calloc((size_t)(*count > 0 ? *count : 1), sizeof(double));
  for (int n = 0; n < *count; ++n)
    ts_double(&ts, &(*values)[n]);
  ts_close(&ts);
  return 0;
}

static int read_local_orbitals(orc_model* m)
{
  char path[4200];
  path_join(path, sizeof(path), m->input_dir, "local_orbitals.in");
  tokstream ts;
  if (!ts_open(&ts, path))
    return fail(m, "Error: cannot open %s", path);
  ts_int(&ts, &m->number_of_local_orbitals);
  int K = m->number_of_local_orbitals > 0 ? m->number_of_local_orbitals : 1;
  m->local_orbitals = (int*)calloc((size_t)K, sizeof(int));
  for (int n = 0; n < m->number_of_local_orbitals; ++n) {
    ts_int(&ts, &m->local_orbitals[n]);
    if (m->local_orbitals[n] < 0 || m->local_orbitals[n] >= m->number_of_atoms) {
      int bad = m->local_orbitals[n];
      ts_close(&ts);
      return fail(m, "Error: wrong local orbital index: %d", bad);
    }
  }
  ts_close(&ts);
  return 0;
}

int orc_model_load(const char* dir, orc_model* m)
{
  memset(m, 0, sizeof(*m));
  m->number_of_random_vectors = 1;
  m->number_of_moments = 1000;
  m->energy_max = 10.0;
  snprintf(m->input_dir, sizeof(m->input_dir), "%s", dir);
  orc_mt_seed(&m->generator, 12345678u);
  if (parse_parameters(m))
    return 1;
  if (m->use_lattice_model ? build_lattice(m) : build_general(m))
    return 1;
  if (m->calculate_spin && (m->number_of_atoms % 2))
    return fail(m, "Error: spin polarization needs an even number of orbitals");
  if (read_counted(m, "energy.in", &m->number_of_energy_points, &m->energy))
    return 1;
  if (m->requires_time &&
      read_counted(m, "time_step.in", &m->number_of_steps_correlation, &m->time_step))
    return 1;
  if (m->calculate_ldos && read_local_orbitals(m))
    return 1;
  return 0;
}

void orc_model_free(orc_model* m)
{
  free(m->neighbor_number);
  free(m->neighbor_list);
  free(m->potential);
  free(m->hopping_real);
  free(m->hopping_imag);
  free(m->xx);
  free(m->energy);
  free(m->time_step);
  free(m->local_orbitals);
  m->neighbor_number = m->neighbor_list = m->local_orbitals = NULL;
  m->potential = m->hopping_real = m->hopping_imag = m->xx = m->energy = m->time_step = NULL;
}

/* src/model.cu:77-109 */
void orc_initialize_state(orc_model* m, int orbital, double* re, double* im)
{
  int N = m->number_of_atoms;
  if (orbital >= 0) {
    for (int n = 0; n < N; ++n) {
      re[n] = 0.0;
      im[n] = 0.0;
    }
    re[orbital] = sqrt(1.0 * N);
  } else if (m->calculate_spin) {
    for (int n = 0; n < N; n += 2) {
      double ph = orc_uniform_real(&m->generator, 0, 2 * ORC_PI);
      re[n] = cos(ph);
      im[n] = sin(ph);
      re[n + 1] = 0.0;
      im[n + 1] = 0.0;
    }
  } else {
    for (int n = 0; n < N; ++n) {
      double ph = orc_uniform_real(&m->generator, 0, 2 * ORC_PI);
      re[n] = cos(ph);
      im[n] = sin(ph);
    }
  }
}
