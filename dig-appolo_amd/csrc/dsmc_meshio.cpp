// Fluent ASCII case (.cas) mesh reading and writing.
//
// The reader follows the reference's readCAS (src/meshImport.cpp:177-576):
// section (2 d) the dimension, (10 ...) nodes, (12 ...) cells, (13 ...) faces;
// counts and indices are lower-case hexadecimal, node and cell ids 1-based; a
// face line lists its nodes then its two cells (c1 = 0 on a boundary), and a
// face zone of type 0 (mixed) starts every line with the node count.  The
// face zone id becomes the face tag (the reference's faceTag, mapped to a role
// by the boundary table: 3 interface, 4 inlet, 5 outlet, 6 wall).  Nodes are
// multiplied by `scale` (the reference's ScaleFactor).  Periodic (18) and zone
// name (45) sections are skipped.  The geometry (normals, areas, volumes,
// cell face lists, quad split tags) is then derived by dsmc_mesh_build, as
// the reference's preprocess does.
//
// The writer emits the same subset: one node zone, one cell zone, and one
// face zone per run of consecutive faces with the same tag.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "dsmc_engine.h"
#include "dsmc_internal.h"

using dsmc::set_error;

namespace {

// hex2dcm (src/meshImport.cpp:553-576): lower-case hex digits, stops at ')',
// skips '('.
long hex(const std::string& s) {
  long v = 0;
  for (char ch : s) {
    if (ch == ')') break;
    if (ch == '(') continue;
    int d;
    if (ch >= '0' && ch <= '9') d = ch - '0';
    else if (ch >= 'a' && ch <= 'f') d = ch - 'a' + 10;
    else if (ch >= 'A' && ch <= 'F') d = ch - 'A' + 10;
    else return -1;
    v = v * 16 + d;
  }
  return v;
}

// The text between the section's header parentheses: "(13 (z f l b t)(" ->
// tokens {z, f, l, b, t}.
std::vector<std::string> header_tokens(const std::string& line) {
  std::vector<std::string> t;
  const size_t a = line.find('(', 1);
  if (a == std::string::npos) return t;
  const size_t b = line.find(')', a);
  std::istringstream is(line.substr(a + 1, (b == std::string::npos ? line.size() : b) - a - 1));
  std::string w;
  while (is >> w) t.push_back(w);
  return t;
}

bool body_opens(const std::string& line) {  // "...)(" ends the header: data follows
  const size_t b = line.find(')', line.find('(', 1));
  return b != std::string::npos && line.find('(', b) != std::string::npos;
}

}  // namespace

extern "C" {

int dsmc_mesh_read_cas(const char* path, double scale, dsmc_mesh** out) {
  if (!path || !out || !(scale > 0.0)) return set_error(DSMC_ERR_ARG, "dsmc_mesh_read_cas: bad arguments");
  std::ifstream in(path);
  if (!in.is_open()) return set_error(DSMC_ERR_ARG, std::string("dsmc_mesh_read_cas: cannot open ") + path);
  long nnode = 0, ncell = 0, nface = 0;
  std::vector<double> xyz;
  std::vector<int32_t> fn, fc, ft, ftag;
  int dim = 3;
  std::string line;
  auto fail = [&](const std::string& why) { return set_error(DSMC_ERR_MESH, "dsmc_mesh_read_cas: " + why); };
  while (std::getline(in, line)) {
    if (line.rfind("(2 ", 0) == 0) {
      dim = std::atoi(line.c_str() + 3);
      if (dim != 3) return fail("only 3-D meshes are supported");
      continue;
    }
    const bool nodes = line.rfind("(10 ", 0) == 0, cells = line.rfind("(12 ", 0) == 0,
               faces = line.rfind("(13 ", 0) == 0;
    if (!nodes && !cells && !faces) continue;
    const std::vector<std::string> t = header_tokens(line);
    if (t.size() < 3 || t[0] == "labels") continue;
    const long zone = hex(t[0]), first = hex(t[1]), last = hex(t[2]);
    if (zone < 0 || first < 0 || last < first - 1) return fail("bad section header: " + line);
    // a zone's records are numbered from 1 (record k goes to index k - 1)
    if (zone != 0 && (first < 1 || last < first)) return fail("bad section range: " + line);
    if (zone == 0) {  // declaration of the totals
      if (nodes) {
        nnode = last;
        xyz.assign((size_t)nnode * 3, 0.0);
      } else if (cells) {
        ncell = last;
      } else {
        nface = last;
        fn.assign((size_t)nface * 4, -1);
        fc.assign((size_t)nface * 2, -1);
        ft.assign((size_t)nface, 0);
        ftag.assign((size_t)nface, 0);
      }
      continue;
    }
    if (cells) {  // cell types: not needed (the faces define the cells)
      if (body_opens(line))
        while (std::getline(in, line) && line.find("))") == std::string::npos) {
        }
      continue;
    }
    if (nodes) {
      if (last > nnode) return fail("node zone beyond the declared node count");
      long k = first - 1;
      while (k < last && std::getline(in, line)) {
        std::istringstream is(line);
        double x, y, z;
        if (!(is >> x >> y >> z)) continue;  // "(" and blank lines
        xyz[3 * k] = x * scale;
        xyz[3 * k + 1] = y * scale;
        xyz[3 * k + 2] = z * scale;
        ++k;
      }
      continue;
    }
    // face zone: (13 (zone first last bc-type face-type)(
    if (t.size() < 5) return fail("bad face header: " + line);
    if (last > nface) return fail("face zone beyond the declared face count");
    const long ztype = hex(t[4]);
    long k = first - 1;
    while (k < last && std::getline(in, line)) {
      std::istringstream is(line);
      std::vector<std::string> w;
      std::string x;
      while (is >> x) w.push_back(x);
      if (w.empty() || w[0] == "(") continue;
      size_t p = 0;
      long nv = ztype;
      if (ztype == 0) nv = hex(w[p++]);
      if (nv != 3 && nv != 4) return fail("only triangle and quad faces are supported");
      if (w.size() < p + nv + 2) return fail("short face line: " + line);
      for (long q = 0; q < nv; ++q) fn[4 * k + q] = (int32_t)(hex(w[p + q]) - 1);
      fc[2 * k] = (int32_t)(hex(w[p + nv]) - 1);
      fc[2 * k + 1] = (int32_t)(hex(w[p + nv + 1]) - 1);
      ft[k] = (int32_t)nv;
      ftag[k] = (int32_t)zone;
      ++k;
    }
    if (k < last) return fail("face zone ended early");
  }
  if (nnode <= 0 || ncell <= 0 || nface <= 0) return fail("missing node, cell or face declarations");
  for (long f = 0; f < nface; ++f)
    if (ft[f] == 0) return fail("a declared face was never listed");
  return dsmc_mesh_build((int32_t)ncell, (int32_t)nface, (int32_t)nnode, xyz.data(), fn.data(), fc.data(), ft.data(),
                         ftag.data(), out);
}

int dsmc_mesh_write_cas(const dsmc_mesh_view* m, const char* path) {
  if (!m || !path) return set_error(DSMC_ERR_ARG, "dsmc_mesh_write_cas: bad arguments");
  std::ofstream out(path);
  if (!out.is_open()) return set_error(DSMC_ERR_ARG, std::string("dsmc_mesh_write_cas: cannot open ") + path);
  out << std::hex;
  out << "(0 \"Fluent ASCII mesh written by the DSMC engine\")\n(2 3)\n";
  out << "(10 (0 1 " << m->nnode << " 0 3))\n";
  out << "(10 (1 1 " << m->nnode << " 1 3)(\n" << std::setprecision(17) << std::scientific;
  for (int n = 0; n < m->nnode; ++n)
    out << m->node_xyz[3 * n] << " " << m->node_xyz[3 * n + 1] << " " << m->node_xyz[3 * n + 2] << "\n";
  out << std::defaultfloat << "))\n";
  out << "(12 (0 1 " << m->ncell << " 0))\n";
  // element types from the cell's faces: 2 tet, 4 hex, 5 pyramid, 6 wedge,
  // 7 polyhedron
  out << "(12 (2 1 " << m->ncell << " 1 0)(\n";
  for (int c = 0; c < m->ncell; ++c) {
    int ntri = 0, nquad = 0;
    for (int j = 0; j < m->cell_nface[c]; ++j) {
      const int f = m->cell_faces[(size_t)c * DSMC_MAX_CELL_FACES + j];
      if (f < 0) continue;
      (m->face_type[f] == 3 ? ntri : nquad) += 1;
    }
    const int type = (ntri == 4 && nquad == 0) ? 2 : (nquad == 6 && ntri == 0) ? 4
                     : (ntri == 4 && nquad == 1) ? 5 : (ntri == 2 && nquad == 3) ? 6 : 7;
    out << type << ((c % 32 == 31 || c + 1 == m->ncell) ? "\n" : " ");
  }
  out << "))\n";
  out << "(13 (0 1 " << m->nface << " 0))\n";
  int f = 0;
  while (f < m->nface) {
    int g = f;
    while (g + 1 < m->nface && m->face_tag[g + 1] == m->face_tag[f]) ++g;
    const int tag = m->face_tag[f];
    // Fluent boundary types: 2 interior, 3 wall, 4 pressure inlet, 5 pressure outlet
    const int bc = tag == DSMC_TAG_WALL ? 3 : tag == DSMC_TAG_INLET ? 4 : tag == DSMC_TAG_OUTLET ? 5 : 2;
    out << "(13 (" << tag << " " << f + 1 << " " << g + 1 << " " << bc << " 0)(\n";
    for (int k = f; k <= g; ++k) {
      const int nv = m->face_type[k];
      out << nv;
      for (int q = 0; q < nv; ++q) out << " " << m->face_nodes[4 * k + q] + 1;
      const int c0 = m->face_cells[2 * k], c1 = m->face_cells[2 * k + 1];
      out << " " << c0 + 1 << " " << ((c1 >= 0 && c1 < m->ncell) ? c1 + 1 : 0) << "\n";
    }
    out << "))\n";
    f = g + 1;
  }
  out.close();
  return out ? DSMC_OK : set_error(DSMC_ERR_ARG, "dsmc_mesh_write_cas: write failed");
}

}  // extern "C"
