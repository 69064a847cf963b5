// sx_io.cpp -- host-side file formats of the drop-in path:
//   * semtex field files (binary nekton format): writeField
//     (src/misc.cpp:106-180), readField (src/misc.cpp:182-222), the Header
//     reader (src/data2df.cpp:454-470) and the binary data layout
//     (AuxField operator<< / >>, src/auxfield.cpp:1021-1126);
//   * session.num numbering files: the enumerate printout
//     (utility/enumerate.cpp:441-552) and BCmgr::buildnum's reader
//     (src/bcmgr.cpp:586-780).
// No GPU code: these run on the host around the solver's C ABI.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "sx_internal.h"

using namespace sx;

namespace {

// "IEEE little-endian" / "IEEE big-endian" (veclib/xbrev.c:16-50)
const char* machine_format ()
{
  const double d = 3.0;
  unsigned char c[8];
  std::memcpy (c, &d, 8);
  if (c[7] == 64 && c[6] == 8) return "IEEE little-endian";
  if (c[0] == 64 && c[1] == 8) return "IEEE big-endian";
  return "unknown";
}

void swap8 (double* v, size_t n)
{
  for (size_t i = 0; i < n; i++) {
    unsigned char* b = reinterpret_cast<unsigned char*> (v + i);
    for (int k = 0; k < 4; k++) std::swap (b[k], b[7 - k]);
  }
}

void copy_str (char* dst, size_t cap, const std::string& s)
{
  std::strncpy (dst, s.c_str (), cap - 1);
  dst[cap - 1] = '\0';
}

// One header as the reference's Header reader takes it: the session and
// creation strings are the first 24 characters of their lines, the
// numbers are read as tokens, the field names as one token and the format
// as the first 25 characters of its line.
bool read_header (std::istream& in, sx_field_header* h)
{
  std::string line;
  auto field = [&] (size_t width, char* dst, size_t cap) {
    if (!std::getline (in, line)) return false;
    std::string v = line.substr (0, std::min (width, line.size ()));
    while (!v.empty () && v.back () == ' ') v.pop_back ();
    copy_str (dst, cap, v);
    return true;
  };
  if (!field (24, h->session, sizeof h->session)) return false;
  if (!field (24, h->created, sizeof h->created)) return false;
  if (!std::getline (in, line)) return false;
  {
    std::istringstream s (line);
    if (!(s >> h->nr >> h->ns >> h->nz >> h->nel)) return false;
  }
  auto number = [&] (auto& v) {
    if (!std::getline (in, line)) return false;
    std::istringstream s (line);
    return static_cast<bool> (s >> v);
  };
  if (!number (h->step) || !number (h->time) || !number (h->dt) || !number (h->kinvis) ||
      !number (h->beta))
    return false;
  if (!std::getline (in, line)) return false;
  {
    std::istringstream s (line);
    std::string f;
    if (!(s >> f)) return false;
    copy_str (h->fields, sizeof h->fields, f);
  }
  return field (25, h->format, sizeof h->format);
}

bool needs_swap (const sx_field_header& h)
{
  const char* m = machine_format ();
  return (std::strstr (m, "big") && std::strstr (h.format, "little")) ||
         (std::strstr (h.format, "big") && std::strstr (m, "little"));
}

int io_error (const std::string& msg)
{
  set_error (msg);
  return SX_ERR_ARG;
}

}  // namespace

extern "C" int sx_field_write (const char* path, const sx_field_header* hdr,
                               const double* const* data, int plane_stride, int append)
{
  if (!path || !hdr || !data || hdr->nr < 2 || hdr->nr != hdr->ns || hdr->nz < 1 ||
      hdr->nel < 1 || plane_stride < hdr->nel * hdr->nr * hdr->ns)
    return io_error ("sx_field_write: bad arguments");
  const size_t nfield = std::strlen (hdr->fields);
  if (nfield < 1) return io_error ("sx_field_write: no fields named");
  for (size_t i = 0; i < nfield; i++)
    if (!data[i]) return io_error ("sx_field_write: missing data for a named field");
  std::ofstream out (path, std::ios::binary | (append ? std::ios::app : std::ios::trunc));
  if (!out) return io_error (std::string ("sx_field_write: cannot open ") + path);
  char s1[256], s2[128];
  std::snprintf (s1, sizeof s1, "%-25s Session\n", hdr->session);
  out << s1;
  if (hdr->created[0]) {
    copy_str (s2, sizeof s2, hdr->created);
  } else {
    const std::time_t tp = std::time (nullptr);
    std::strftime (s2, 25, "%a %b %d %H:%M:%S %Y", std::localtime (&tp));
  }
  std::snprintf (s1, sizeof s1, "%-25s Created\n", s2);
  out << s1;
  // AuxField::describe (src/auxfield.cpp:1143-1154): "np np nz nel"
  std::snprintf (s2, sizeof s2, "%d %d %d %d", hdr->nr, hdr->ns, hdr->nz, hdr->nel);
  std::snprintf (s1, sizeof s1, "%-25s Nr, Ns, Nz, Elements\n", s2);
  out << s1;
  std::snprintf (s1, sizeof s1, "%-25d Step\n", hdr->step);
  out << s1;
  std::snprintf (s1, sizeof s1, "%-25.6g Time\n", hdr->time);
  out << s1;
  std::snprintf (s1, sizeof s1, "%-25.6g Time step\n", hdr->dt);
  out << s1;
  std::snprintf (s1, sizeof s1, "%-25.6g Kinvis\n", hdr->kinvis);
  out << s1;
  std::snprintf (s1, sizeof s1, "%-25.6g Beta\n", hdr->beta);
  out << s1;
  std::snprintf (s1, sizeof s1, "%-25s Fields written\n", hdr->fields);
  out << s1;
  std::snprintf (s2, sizeof s2, "binary %s", machine_format ());
  std::snprintf (s1, sizeof s1, "%-25s Format\n", s2);
  out << s1;
  // each field: nz planes of nel np^2 values (Geometry::nPlane, no padding)
  const size_t nplane = (size_t) hdr->nel * hdr->nr * hdr->ns;
  for (size_t f = 0; f < nfield; f++)
    for (int k = 0; k < hdr->nz; k++)
      out.write (reinterpret_cast<const char*> (data[f] + (size_t) k * plane_stride),
                 (std::streamsize) (sizeof (double) * nplane));
  out.flush ();
  if (!out) return io_error ("sx_field_write: failed writing field file");
  return SX_OK;
}

extern "C" int sx_field_read_header (const char* path, int dump, sx_field_header* hdr)
{
  if (!path || !hdr || dump < 0) return io_error ("sx_field_read_header: bad arguments");
  std::ifstream in (path, std::ios::binary);
  if (!in) return io_error (std::string ("sx_field_read_header: cannot open ") + path);
  for (int d = 0;; d++) {
    std::memset (hdr, 0, sizeof *hdr);
    if (!read_header (in, hdr))
      return io_error ("sx_field_read_header: no dump " + std::to_string (dump) + " in " + path);
    if (d == dump) return SX_OK;
    const size_t bytes = sizeof (double) * std::strlen (hdr->fields) * (size_t) hdr->nz *
                         hdr->nel * hdr->nr * hdr->ns;
    in.seekg ((std::streamoff) bytes, std::ios::cur);
  }
}

extern "C" int sx_field_read (const char* path, int dump, char name, double* data,
                              int plane_stride)
{
  if (!path || !data || dump < 0) return io_error ("sx_field_read: bad arguments");
  std::ifstream in (path, std::ios::binary);
  if (!in) return io_error (std::string ("sx_field_read: cannot open ") + path);
  sx_field_header h;
  for (int d = 0;; d++) {
    std::memset (&h, 0, sizeof h);
    if (!read_header (in, &h))
      return io_error ("sx_field_read: no dump " + std::to_string (dump) + " in " + path);
    const size_t nplane = (size_t) h.nel * h.nr * h.ns;
    const size_t per = sizeof (double) * (size_t) h.nz * nplane;
    if (d < dump) {
      in.seekg ((std::streamoff) (per * std::strlen (h.fields)), std::ios::cur);
      continue;
    }
    if (plane_stride < (int) nplane) return io_error ("sx_field_read: plane stride too small");
    if (!std::strstr (h.format, "binary"))
      return io_error ("sx_field_read: field file not in binary format");
    const char* at = std::strchr (h.fields, name);
    if (!at) return io_error (std::string ("sx_field_read: no field '") + name + "' in dump");
    in.seekg ((std::streamoff) (per * (size_t) (at - h.fields)), std::ios::cur);
    const bool swap = needs_swap (h);
    for (int k = 0; k < h.nz; k++) {
      double* p = data + (size_t) k * plane_stride;
      in.read (reinterpret_cast<char*> (p), (std::streamsize) (sizeof (double) * nplane));
      if (!in) return io_error ("sx_field_read: unable to read binary input");
      if (swap) swap8 (p, nplane);
      // the plane padding is zeroed (AuxField operator>>)
      for (size_t i = nplane; i < (size_t) plane_stride; i++) p[i] = 0.0;
    }
    return SX_OK;
  }
}

// ---------------------------------------------------------------------------
// session.num
// ---------------------------------------------------------------------------

extern "C" int sx_num_write (const char* path, const char* fields, int nset,
                             const char* const* set_fields, int nel, int np,
                             const int* const* btog, const int* const* bmask,
                             const int* nglobal, const int* nsolve, const int* optlev,
                             const int* nband)
{
  if (!path || !fields || nset < 1 || !set_fields || nel < 1 || np < 2 || !btog || !bmask ||
      !nglobal || !nsolve || !optlev || !nband)
    return io_error ("sx_num_write: bad arguments");
  std::ofstream o (path, std::ios::trunc);
  if (!o) return io_error (std::string ("sx_num_write: cannot open ") + path);
  const int next = 4 * (np - 1), nint = (np - 2) * (np - 2);
  const int nbndry = nel * next, ntotal = nel * np * np;
  auto sep = [&] () {
    o << "# ----------------";
    for (int j = 0; j < nset; j++) o << "  -------------";
    o << "\n";
  };
  auto row = [&] (const char* label, auto value) {
    o << label;
    for (int j = 0; j < nset; j++) o << std::setw (15) << value (j);
    o << "\n";
  };
  o << "# FIELDS         :  " << fields << "\n";
  sep ();
  o << "# " << nset << " NUMBER SETS  :";
  for (int j = 0; j < nset; j++) o << std::setw (15) << set_fields[j];
  o << "\n";
  row ("# NEL            :", [&] (int) { return nel; });
  row ("# NP_MAX         :", [&] (int) { return np; });
  row ("# NEXT_MAX       :", [&] (int) { return next; });
  row ("# NINT_MAX       :", [&] (int) { return nint; });
  row ("# NTOTAL         :", [&] (int) { return ntotal; });
  row ("# NBOUNDARY      :", [&] (int) { return nbndry; });
  row ("# NGLOBAL        :", [&] (int j) { return nglobal[j]; });
  row ("# NSOLVE         :", [&] (int j) { return nsolve[j]; });
  row ("# OPTIMIZATION   :", [&] (int j) { return optlev[j]; });
  row ("# BANDWIDTH      :", [&] (int j) { return nband[j]; });
  sep ();
  o << "# elmt  side offst";
  for (int j = 0; j < nset; j++) o << "     bmap  mask";
  o << "\n";
  const int nedge = np - 1;
  for (int i = 0, k = 1; k <= nel; k++)
    for (int side = 1; side <= 4; side++)
      for (int soff = 0; soff < nedge; soff++, i++) {
        o << std::setw (6) << k << std::setw (6) << side << std::setw (6) << soff;
        for (int j = 0; j < nset; j++)
          o << std::setw (9) << btog[j][i] << std::setw (6) << bmask[j][i];
        o << "\n";
      }
  o.flush ();
  if (!o) return io_error ("sx_num_write: failed writing");
  return SX_OK;
}

extern "C" int sx_num_read (const char* path, int nel, int np, int max_sets, int* nset,
                            char* set_fields, int field_cap, int* btog, int* bmask,
                            int* nglobal, int* nsolve, int* optlev, int* nband)
{
  if (!path || nel < 1 || np < 2 || max_sets < 1 || !nset || !set_fields || field_cap < 2 ||
      !btog || !bmask || !nglobal || !nsolve || !optlev || !nband)
    return io_error ("sx_num_read: bad arguments");
  std::ifstream in (path);
  if (!in) return io_error (std::string ("sx_num_read: cannot open ") + path);
  std::string buf, line;
  in >> buf >> buf >> buf >> buf;            // "# FIELDS : uvwp"
  std::getline (in, line);
  std::getline (in, line);                   // separator
  int n = 0;
  in >> buf >> n >> buf >> buf >> buf;       // "# 2 NUMBER SETS :"
  if (!in || n < 1 || n > max_sets) return io_error ("sx_num_read: bad NUMBER SETS line");
  *nset = n;
  for (int i = 0; i < n; i++) {
    in >> buf;
    if ((int) buf.size () >= field_cap) return io_error ("sx_num_read: field names too long");
    copy_str (set_fields + (size_t) i * field_cap, (size_t) field_cap, buf);
  }
  const int next = 4 * (np - 1), ntotal = nel * np * np, nbndry = nel * next;
  auto expect = [&] (const char* key, auto check) {
    in >> buf >> buf;
    if (buf != key) return io_error (std::string ("sx_num_read: expected \"") + key +
                                     "\", found " + buf);
    in >> buf;                                // ":"
    for (int j = 0; j < n; j++) {
      int v;
      in >> v;
      if (!in) return io_error (std::string ("sx_num_read: bad ") + key + " line");
      if (int rc = check (j, v)) return rc;
    }
    return (int) SX_OK;
  };
  auto must = [&] (const char* what, int want) {
    return [=] (int, int v) {
      return v == want ? (int) SX_OK
                       : io_error (std::string ("sx_num_read: mismatch in ") + what + ": " +
                                   std::to_string (v) + " vs. " + std::to_string (want));
    };
  };
  auto any = [] (int, int) { return (int) SX_OK; };
  int rc;
  if ((rc = expect ("NEL", must ("number of elements", nel))) ||
      (rc = expect ("NP_MAX", must ("element order", np))) ||
      (rc = expect ("NEXT_MAX", any)) || (rc = expect ("NINT_MAX", any)) ||
      (rc = expect ("NTOTAL", must ("Field storage requirements", ntotal))) ||
      (rc = expect ("NBOUNDARY", must ("number of boundary nodes", nbndry))) ||
      (rc = expect ("NGLOBAL", [&] (int j, int v) { nglobal[j] = v; return (int) SX_OK; })) ||
      (rc = expect ("NSOLVE", [&] (int j, int v) { nsolve[j] = v; return (int) SX_OK; })) ||
      (rc = expect ("OPTIMIZATION", [&] (int j, int v) { optlev[j] = v; return (int) SX_OK; })) ||
      (rc = expect ("BANDWIDTH", [&] (int j, int v) { nband[j] = v; return (int) SX_OK; })))
    return rc;
  std::getline (in, line);
  std::getline (in, line);                   // separator
  std::getline (in, line);                   // column titles
  for (int i = 0; i < nbndry; i++) {
    int e, s, o;
    in >> e >> s >> o;
    for (int j = 0; j < n; j++) in >> btog[(size_t) j * nbndry + i] >> bmask[(size_t) j * nbndry + i];
    if (!in) return io_error ("sx_num_read: failed reading to end of node-number file");
  }
  return SX_OK;
}
