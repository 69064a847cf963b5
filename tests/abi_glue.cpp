// abi_glue.cpp -- a C++ caller of libsemtex_amd that makes the calls the
// drop-in glue in semtex's drivers makes (the hooks in ModalMatrixSys,
// Field::solve and integrate(); dns/drive.cpp:54-55, src/field.h:27,
// src/matrix.h:61-66), with plain host arrays standing in for semtex's
// Element / NumberSys / BoundarySys objects.  Built by the test suite with
// g++ against include/semtex_amd.h and linked to libsemtex_amd.so.
//
//   abi_glue <case-dir>
//
// <case-dir> holds params.txt ("key value" lines) and raw little-endian
// arrays written by the test: for each numbering sys in {u, p}:
// <sys>_x, <sys>_y (double [nel][np*np]), <sys>_btog, <sys>_bmask
// (int [nel][next]), <sys>_emask (int [nel]), <sys>_bnd (int [nbound][3]:
// element, side, kind); helm_f, helm_lines (double [nz][stride],
// [nz][nline]); dns_u0 (double [3][nz][stride], physical space).
// Writes helm_u (the Field::solve result) and dns_out (u, v, w, p in
// Fourier space after `dns_steps` steps, [4][nz][stride]).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "semtex_amd.h"

namespace {

std::string dir;
std::map<std::string, double> par;

template <class T> std::vector<T> load (const std::string& name, size_t n)
{
  std::vector<T> v (n);
  FILE* f = std::fopen ((dir + "/" + name).c_str (), "rb");
  if (!f || std::fread (v.data (), sizeof (T), n, f) != n) {
    std::fprintf (stderr, "abi_glue: cannot read %s (%zu items)\n", name.c_str (), n);
    std::exit (2);
  }
  std::fclose (f);
  return v;
}

void save (const std::string& name, const std::vector<double>& v)
{
  FILE* f = std::fopen ((dir + "/" + name).c_str (), "wb");
  if (!f || std::fwrite (v.data (), sizeof (double), v.size (), f) != v.size ()) {
    std::fprintf (stderr, "abi_glue: cannot write %s\n", name.c_str ());
    std::exit (2);
  }
  std::fclose (f);
}

int ip (const std::string& k) { return (int) par.at (k); }

// The semtex-side objects the glue reads through accessors: element nodes
// (Element::_xmesh/_ymesh), a numbering (NumberSys::btog/bmask/emask/
// nGlobal/nSolve/nBand) and boundary segments (BoundarySys::BCs).
struct Elements { int np, nel; std::vector<double> x, y; };
struct NumberSys { std::vector<int> btog, bmask, emask; int nglobal, nsolve, nband; };
struct BoundarySys { std::vector<int> elmt, side, kind; const NumberSys* N; };

// Emulates semtex's message(routine, text, ERROR): print and exit.
void message (const char* routine, const char* text)
{
  std::fprintf (stderr, "%s: %s\n", routine, text);
  std::exit (EXIT_FAILURE);
}

// One GPU mesh per numbering system: velocity and pressure have different
// essential sets, so the cache is keyed by the NumberSys.
sx_mesh_t gpuMesh (const Elements& E, const BoundarySys* B, const NumberSys* N)
{
  static std::map<const NumberSys*, sx_mesh_t> cache;
  auto it = cache.find (N);
  if (it != cache.end ()) return it->second;
  sx_mesh_t M = nullptr;
  const int rc = sx_mesh_create (E.np, E.nel, 0, E.x.data (), E.y.data (), N->btog.data (),
                                 N->bmask.data (), N->emask.data (), N->nglobal, N->nsolve,
                                 N->nband, (int) B->elmt.size (), B->elmt.data (),
                                 B->side.data (), B->kind.data (), &M);
  if (rc) message ("gpuMesh", sx_last_error ());
  cache[N] = M;
  return M;
}

void read_sys (const std::string& s, const Elements& E, NumberSys& N, BoundarySys& B)
{
  const int next = 4 * (E.np - 1), nb = ip (s + "_nbound");
  N.btog = load<int> (s + "_btog", (size_t) E.nel * next);
  N.bmask = load<int> (s + "_bmask", (size_t) E.nel * next);
  N.emask = load<int> (s + "_emask", (size_t) E.nel);
  N.nglobal = ip (s + "_nglobal");
  N.nsolve = ip (s + "_nsolve");
  N.nband = ip (s + "_nband");
  const std::vector<int> b = load<int> (s + "_bnd", (size_t) 3 * nb);
  for (int i = 0; i < nb; i++) {
    B.elmt.push_back (b[3 * i]);
    B.side.push_back (b[3 * i + 1]);
    B.kind.push_back (b[3 * i + 2]);
  }
  B.N = &N;
}

}  // namespace

int main (int argc, char** argv)
{
  if (argc != 2) { std::fprintf (stderr, "usage: abi_glue <case-dir>\n"); return 2; }
  dir = argv[1];
  {
    FILE* f = std::fopen ((dir + "/params.txt").c_str (), "r");
    if (!f) { std::fprintf (stderr, "abi_glue: no params.txt\n"); return 2; }
    char key[128];
    double v;
    while (std::fscanf (f, "%127s %lf", key, &v) == 2) par[key] = v;
    std::fclose (f);
  }
  Elements E;
  E.np = ip ("np");
  E.nel = ip ("nel");
  const size_t npl = (size_t) E.nel * E.np * E.np;
  E.x = load<double> ("x", npl);
  E.y = load<double> ("y", npl);
  NumberSys Nu, Np;
  BoundarySys Bu, Bp;
  read_sys ("u", E, Nu, Bu);
  read_sys ("p", E, Np, Bp);
  const int nz = ip ("nz"), stride = ip ("stride"), nline = ip ("u_nbound") * E.np;

  // Hook 1 + 2: ModalMatrixSys (DIRECT) and Field::solve with host data,
  // on the velocity numbering (Dirichlet walls).
  {
    sx_msys_t S = nullptr;
    sx_mesh_t M = gpuMesh (E, &Bu, &Nu);
    if (sx_msys_create (M, par.at ("lambda2"), par.at ("beta"), 0, nz / 2, SX_FP64, nullptr, &S))
      message ("ModalMatrixSys", sx_last_error ());
    std::vector<double> f = load<double> ("helm_f", (size_t) nz * stride);
    std::vector<double> lines = load<double> ("helm_lines", (size_t) nz * nline);
    std::vector<double> u ((size_t) nz * stride, 0.0);
    if (sx_field_solve_host (S, nz, stride, f.data (), u.data (), lines.data ()))
      message ("Field::solve", sx_last_error ());
    save ("helm_u", u);
    sx_msys_destroy (S);
  }

  // Hook 3: integrate() on the device, velocity and pressure numberings
  // from one cache; the fields come back to the host at the end.
  {
    sx_mesh_t Mu = gpuMesh (E, &Bu, &Nu);
    sx_mesh_t Mp = gpuMesh (E, &Bp, &Np);
    if (Mu == Mp) message ("integrate", "velocity and pressure share a GPU mesh");
    const double cf[3] = {par.at ("cf_x"), 0.0, 0.0};
    sx_dns_t G = nullptr;
    if (sx_dns_create (Mu, Mp, nz, par.at ("beta"), par.at ("kinvis"), par.at ("dt"), cf,
                       SX_SKEW, 2, nullptr, &G))
      message ("integrate", sx_last_error ());
    const int S = sx_dns_plane_stride (G);
    if (S != stride) message ("integrate", "plane stride differs from the field's");
    const size_t nf = (size_t) nz * S;
    std::vector<double> u0 = load<double> ("dns_u0", 3 * nf);
    double* d_u = nullptr;
    if (hipMalloc (&d_u, sizeof (double) * 3 * nf) != hipSuccess ||
        hipMemcpy (d_u, u0.data (), sizeof (double) * 3 * nf, hipMemcpyHostToDevice) != hipSuccess)
      message ("integrate", "device copy of the initial field failed");
    if (sx_dns_set_state (G, d_u, nullptr)) message ("integrate", sx_last_error ());
    const int nsteps = ip ("dns_steps");
    for (int step = 0; step < nsteps; step++)
      if (sx_dns_step (G, 1, nullptr)) message ("integrate", sx_last_error ());
    std::vector<double> out (4 * nf);
    for (int i = 0; i < 4; i++) {
      int st = 0;
      if (sx_dns_get (G, i, d_u, &st, nullptr) ||
          hipMemcpy (&out[i * nf], d_u, sizeof (double) * nf, hipMemcpyDeviceToHost) != hipSuccess)
        message ("integrate", "field copy failed");
      if (st != nsteps) message ("integrate", "step count differs");
    }
    save ("dns_out", out);
    (void) hipFree (d_u);
    sx_dns_destroy (G);
  }
  std::printf ("abi_glue: ok\n");
  return 0;
}
