// molio.cpp -- ligand input on the host: .mol2 -> topology -> .pqrs, and the .pqrs reader.
//
// The reference prepares ligands with mol2pqrs and reads them back in lds:
//   Mol2 parsing, lone pairs removed            src/mol2.ml:99-113, :150-323
//   mol2 type -> element, vdW radii             src/ptable.ml:43-53, :169-225
//   topological distances (BFS, unit weights)   src/mol_graph.ml:53-95
//   rotatable bonds: single (types 1, am), both ends of degree > 1, not in a ring
//                                               src/mol_graph.ml:120-160
//   rotation group: the smaller side of the bond (tie: the side without the first atom)
//                                               src/mol_graph.ml:162-192
//   .pqrs block "n:nrb:name", atom lines "%g %g %g %g %g symb", "beg-end= b0 .. bn-1"
//   lines, distance matrix rows                 src/mol2pqrs.ml:9-43
//   .pqrs reader (axis from the fixed atom to the moving one, group without the axis
//   end, column-major distances)                src/pqrs.ml:17-87, src/mol.ml:368-402
// A .mol2 file is converted to .pqrs text and read back, as the reference's two tools
// do, so both inputs give identical molecules (coordinates rounded by %g included).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "mmo_ff.h"

namespace mmo {
int host_fail(int code, const std::string& msg);   // api.hip: sets mmo_last_error(NULL)
}

namespace {

struct InputError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// Ptable.anum_of_mol2_type (src/ptable.ml:169-225)
const std::map<std::string, int> kMol2TypeAnum = {
    {"H", 1},      {"H.t3p", 1},  {"H.spc", 1},  {"Li", 3},     {"C.3", 6},    {"C.2", 6},
    {"C.1", 6},    {"C.ar", 6},   {"C.cat", 6},  {"N.3", 7},    {"N.2", 7},    {"N.1", 7},
    {"N.ar", 7},   {"N.am", 7},   {"N.pl3", 7},  {"N.4", 7},    {"O.3", 8},    {"O.2", 8},
    {"O.co2", 8},  {"O.spc", 8},  {"O.t3p", 8},  {"F", 9},      {"Na", 11},    {"Mg", 12},
    {"Al", 13},    {"Si", 14},    {"P.3", 15},   {"S.3", 16},   {"S.2", 16},   {"S.O", 16},
    {"S.o", 16},   {"S.O2", 16},  {"S.o2", 16},  {"Cl", 17},    {"K", 19},     {"Ca", 20},
    {"Cr.th", 24}, {"Cr.oh", 24}, {"Mn", 25},    {"Fe", 26},    {"Co.oh", 27}, {"Cu", 29},
    {"Zn", 30},    {"Se", 34},    {"Br", 35},    {"Mo", 42},    {"Sn", 50},    {"I", 53}};
// Ptable.vdW_radii (src/ptable.ml:43-53); other elements have none (NaN)
const std::map<int, double> kVdwRadius = {{1, 1.2},  {6, 1.7},  {7, 1.6},  {8, 1.55},
                                          {9, 1.5},  {12, 2.2}, {15, 1.95}, {16, 1.8},
                                          {17, 1.8}, {35, 1.9}, {53, 2.1}};
// Mol2.bond_order (src/mol2.ml:140-150)
const std::map<std::string, double> kBondOrder = {
    {"1", 1.0}, {"2", 2.0}, {"3", 3.0}, {"ar", 1.5}, {"am", 1.0}};
// element symbols of .pqrs files (src/ptable.ml:86-107)
const std::map<std::string, int> kSymbolAnum = {{"C", 6},  {"H", 1},  {"N", 7},   {"O", 8},
                                                {"P", 15}, {"S", 16}, {"F", 9},   {"Cl", 17},
                                                {"Br", 35}, {"I", 53}, {"Mg", 12}};

std::vector<std::string> split_ws(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && std::isspace((unsigned char)s[i])) ++i;
    size_t j = i;
    while (j < s.size() && !std::isspace((unsigned char)s[j])) ++j;
    if (j > i) out.emplace_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

std::string strip(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::vector<std::string> lines_of(const char* text, size_t len) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < len) {
    size_t j = i;
    while (j < len && text[j] != '\n') ++j;
    std::string l(text + i, j - i);
    if (!l.empty() && l.back() == '\r') l.pop_back();
    out.push_back(std::move(l));
    i = j + 1;
  }
  return out;
}

long to_int(const std::string& s, const char* what) {
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(s.c_str(), &end, 10);
  if (s.empty() || *end != '\0' || errno) throw InputError(std::string("bad ") + what + ": " + s);
  return v;
}

double to_double(const std::string& s, const char* what) {
  char* end = nullptr;
  const double v = std::strtod(s.c_str(), &end);
  if (s.empty() || *end != '\0') throw InputError(std::string("bad ") + what + ": " + s);
  return v;
}

std::string fmt_g(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%g", v);
  return buf;
}

// ------------------------------------------------------------------------ .mol2
struct M2Atom {
  double x, y, z, q, r;
  int anum;
  std::string symb;
};
struct M2Bond {
  int s, d;
  double order;
};
struct M2Mol {
  std::string name;
  std::vector<M2Atom> atoms;
  std::vector<M2Bond> bonds;
};

M2Mol parse_mol2_block(const std::vector<std::string>& b) {
  if (b.size() < 3) throw InputError("truncated @<TRIPOS>MOLECULE block");
  M2Mol m;
  m.name = strip(b[1]);
  const auto counts = split_ws(b[2]);
  if (counts.size() < 2) throw InputError(m.name + ": no atom/bond counts");
  const long n_atoms = to_int(counts[0], "atom count"), n_bonds = to_int(counts[1], "bond count");
  size_t i = 3;
  while (i < b.size() && strip(b[i]) != "@<TRIPOS>ATOM") ++i;
  if (i == b.size()) throw InputError(m.name + ": no @<TRIPOS>ATOM section");
  ++i;
  std::vector<bool> lone;
  std::vector<M2Atom> atoms;
  for (long k = 0; k < n_atoms; ++k, ++i) {
    if (i >= b.size()) throw InputError(m.name + ": truncated atom section");
    const auto f = split_ws(b[i]);
    if (f.size() < 6) throw InputError(m.name + ": bad atom line: " + b[i]);
    const std::string& typ = f[5];
    const double q = f.size() > 8 ? to_double(f[8], "charge") : 0.0;
    if (typ == "LP") {   // lone pair: dropped with its bonds
      lone.push_back(true);
      atoms.push_back(M2Atom{0, 0, 0, 0, 0, 0, ""});
      continue;
    }
    const auto t = kMol2TypeAnum.find(typ);
    if (t == kMol2TypeAnum.end()) throw InputError(m.name + ": unsupported mol2 atom type " + typ);
    const auto rv = kVdwRadius.find(t->second);
    M2Atom a{to_double(f[2], "x"), to_double(f[3], "y"), to_double(f[4], "z"), q,
             rv == kVdwRadius.end() ? std::nan("") : rv->second, t->second,
             typ.substr(0, typ.find('.'))};
    lone.push_back(false);
    atoms.push_back(a);
  }
  while (i < b.size() && strip(b[i]) != "@<TRIPOS>BOND") ++i;
  if (i == b.size() && n_bonds > 0) throw InputError(m.name + ": no @<TRIPOS>BOND section");
  ++i;
  std::vector<M2Bond> bonds;
  for (long k = 0; k < n_bonds; ++k, ++i) {
    if (i >= b.size()) throw InputError(m.name + ": truncated bond section");
    const auto f = split_ws(b[i]);
    if (f.size() < 4) throw InputError(m.name + ": bad bond line: " + b[i]);
    const auto o = kBondOrder.find(f[3]);
    if (o == kBondOrder.end()) throw InputError(m.name + ": unsupported bond type " + f[3]);
    const long s = to_int(f[1], "bond atom") - 1, d = to_int(f[2], "bond atom") - 1;
    if (s < 0 || d < 0 || s >= n_atoms || d >= n_atoms)
      throw InputError(m.name + ": bond atom out of range: " + b[i]);
    bonds.push_back(M2Bond{(int)s, (int)d, o->second});
  }
  // Mol2.remove_lone_pairs: keep the other atoms, renumber the bonds between them
  std::vector<int> renum(atoms.size(), -1);
  for (size_t k = 0; k < atoms.size(); ++k)
    if (!lone[k]) {
      renum[k] = (int)m.atoms.size();
      m.atoms.push_back(atoms[k]);
    }
  for (const M2Bond& bd : bonds)
    if (renum[bd.s] >= 0 && renum[bd.d] >= 0) m.bonds.push_back({renum[bd.s], renum[bd.d], bd.order});
  return m;
}

std::vector<M2Mol> parse_mol2(const std::vector<std::string>& lines) {
  std::vector<M2Mol> out;
  std::vector<std::string> block;
  bool in = false;
  for (const std::string& l : lines) {
    if (strip(l) == "@<TRIPOS>MOLECULE") {
      if (in) out.push_back(parse_mol2_block(block));
      block.assign(1, l);
      in = true;
    } else if (in) {
      block.push_back(l);
    }
  }
  if (in) out.push_back(parse_mol2_block(block));
  return out;
}

// ------------------------------------------------------------------ topology
using Adj = std::vector<std::vector<int>>;

Adj adjacency(const M2Mol& m, int skip) {
  Adj adj(m.atoms.size());
  for (size_t k = 0; k < m.bonds.size(); ++k)
    if ((int)k != skip) {
      adj[m.bonds[k].s].push_back(m.bonds[k].d);
      adj[m.bonds[k].d].push_back(m.bonds[k].s);
    }
  return adj;
}

std::vector<int> bfs(const Adj& adj, int src) {
  std::vector<int> dist(adj.size(), -1);
  std::deque<int> q;
  dist[src] = 0;
  q.push_back(src);
  while (!q.empty()) {
    const int u = q.front();
    q.pop_front();
    for (int v : adj[u])
      if (dist[v] < 0) {
        dist[v] = dist[u] + 1;
        q.push_back(v);
      }
  }
  return dist;
}

// the .pqrs block Mol2pqrs.process_ligand writes for one molecule
std::string pqrs_block(const M2Mol& m) {
  const int n = (int)m.atoms.size();
  const Adj adj = adjacency(m, -1);
  std::vector<std::vector<int>> dm(n);
  for (int i = 0; i < n; ++i) {
    dm[i] = bfs(adj, i);
    for (int v : dm[i])
      if (v < 0) throw InputError("disconnected atom in " + m.name);
  }
  std::vector<int> deg(n, 0);
  for (const M2Bond& b : m.bonds) {
    ++deg[b.s];
    ++deg[b.d];
  }
  std::vector<int> rot;
  for (size_t k = 0; k < m.bonds.size(); ++k) {
    const M2Bond& b = m.bonds[k];
    if (b.order == 1.0 && deg[b.s] > 1 && deg[b.d] > 1 && bfs(adjacency(m, (int)k), b.s)[b.d] < 0)
      rot.push_back((int)k);
  }
  std::string out = std::to_string(n) + ":" + std::to_string(rot.size()) + ":" + m.name + "\n";
  for (const M2Atom& a : m.atoms)
    out += fmt_g(a.x) + " " + fmt_g(a.y) + " " + fmt_g(a.z) + " " + fmt_g(a.q) + " " + fmt_g(a.r) +
           " " + a.symb + "\n";
  for (int k : rot) {
    const M2Bond& b = m.bonds[k];
    const std::vector<int> side_d = bfs(adjacency(m, k), b.d);
    int n_d = 0;
    for (int v : side_d) n_d += v >= 0;
    const bool take_d = n_d <= n - n_d;   // the smaller side moves
    out += std::to_string(b.s) + "-" + std::to_string(b.d) + "=";
    for (int i = 0; i < n; ++i) out += ((side_d[i] >= 0) == take_d) ? " 1" : " 0";
    out += "\n";
  }
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) {
      if (j) out += " ";
      out += std::to_string(dm[i][j]);
    }
    out += "\n";
  }
  return out;
}

// ----------------------------------------------------------------------- .pqrs
struct Molecule {
  std::string name;
  bool ligand = true;
  std::vector<double> x, y, z, q, r;
  std::vector<int32_t> anum;
  std::vector<int32_t> dists;            // i + j n
  std::vector<int32_t> rb_left, rb_right;
  std::vector<int32_t> rg_off{0}, rg_atoms;
};

void parse_header(const std::string& line, long& n, long& nrb, std::string& name) {
  std::vector<std::string> parts;
  const std::string s = strip(line);
  size_t i = 0;
  while (true) {
    const size_t j = s.find(':', i);
    parts.push_back(s.substr(i, j == std::string::npos ? std::string::npos : j - i));
    if (j == std::string::npos) break;
    i = j + 1;
  }
  auto join = [&](size_t from) {
    std::string r;
    for (size_t k = from; k < parts.size(); ++k) r += (k > from ? ":" : "") + parts[k];
    return r;
  };
  if (parts.size() >= 3) {
    try {
      n = to_int(strip(parts[0]), "atom count");
      nrb = to_int(strip(parts[1]), "rot-bond count");
      name = join(2);
      return;
    } catch (const InputError&) {
    }
  }
  if (parts.size() < 2) throw InputError("bad .pqrs header: " + line);
  n = to_int(strip(parts[0]), "atom count");
  nrb = -1;
  name = join(1);
}

Molecule read_one(const std::vector<std::string>& lines, size_t& pos, bool ligand) {
  Molecule m;
  m.ligand = ligand;
  long n = 0, nrb = -1;
  parse_header(lines[pos++], n, nrb, m.name);
  if (n < 0) throw InputError("negative atom count in " + m.name);
  for (long k = 0; k < n; ++k) {
    if (pos >= lines.size()) throw InputError("truncated atoms of " + m.name);
    const auto f = split_ws(lines[pos++]);
    if (f.size() < 6) throw InputError("bad atom line in " + m.name);
    m.x.push_back(to_double(f[0], "x"));
    m.y.push_back(to_double(f[1], "y"));
    m.z.push_back(to_double(f[2], "z"));
    m.q.push_back(to_double(f[3], "charge"));
    m.r.push_back(to_double(f[4], "radius"));
    const auto e = kSymbolAnum.find(f[5]);
    if (e == kSymbolAnum.end()) throw InputError("unsupported element: " + f[5]);
    m.anum.push_back(e->second);
  }
  if (!ligand) return m;
  for (long k = 0; k < std::max(nrb, 0L); ++k) {
    if (pos >= lines.size()) throw InputError("truncated rot-bonds of " + m.name);
    const std::string& l = lines[pos++];
    const size_t eq = l.find('='), dash = l.find('-');
    if (eq == std::string::npos || dash == std::string::npos || dash > eq)
      throw InputError("bad rot-bond line: " + l);
    const long beg = to_int(strip(l.substr(0, dash)), "rot-bond atom");
    const long end = to_int(strip(l.substr(dash + 1, eq - dash - 1)), "rot-bond atom");
    const auto bits = split_ws(l.substr(eq + 1));
    if ((long)bits.size() != n || beg < 0 || end < 0 || beg >= n || end >= n)
      throw InputError("bad rot-bond line: " + l);
    std::vector<long> bv(n);
    for (long i = 0; i < n; ++i) bv[i] = to_int(bits[i], "rot-group flag");
    if ((bv[beg] != 0) == (bv[end] != 0)) throw InputError("bad rot-bond line: " + l);
    // axis from the fixed atom to the moving one
    const long left = bv[end] ? beg : end, right = bv[end] ? end : beg;
    m.rb_left.push_back((int32_t)left);
    m.rb_right.push_back((int32_t)right);
    for (long i = 0; i < n; ++i)
      if (bv[i] && i != right) m.rg_atoms.push_back((int32_t)i);
    m.rg_off.push_back((int32_t)m.rg_atoms.size());
  }
  m.dists.assign((size_t)n * n, 0);
  for (long i = 0; i < n; ++i) {
    if (pos >= lines.size()) throw InputError("truncated distance matrix of " + m.name);
    const auto row = split_ws(lines[pos++]);
    if ((long)row.size() != n) throw InputError("bad distance matrix row in " + m.name);
    for (long j = 0; j < n; ++j) m.dists[(size_t)i + (size_t)j * n] = (int32_t)to_int(row[j], "distance");
    if (m.dists[(size_t)i + (size_t)i * n] != 0) throw InputError("bad distance matrix row in " + m.name);
  }
  return m;
}

std::vector<Molecule> parse_pqrs(const char* text, size_t len, bool ligand) {
  std::vector<std::string> lines;
  for (std::string& l : lines_of(text, len))
    if (!strip(l).empty()) lines.push_back(std::move(l));
  std::vector<Molecule> out;
  std::set<std::string> seen;
  size_t pos = 0;
  while (pos < lines.size()) {
    out.push_back(read_one(lines, pos, ligand));
    if (!seen.insert(out.back().name).second)
      throw InputError("duplicate molecule name " + out.back().name);
    if (!ligand) break;   // a receptor file holds one molecule
  }
  return out;
}

}  // namespace

struct mmo_molset {
  std::vector<Molecule> mols;
  std::string pqrs;
};

extern "C" {

int mmo_molset_from_mol2(const char* text, size_t len, mmo_molset** out) {
  if (!out || (!text && len)) return mmo::host_fail(MMO_ERR_BAD_ARG, "NULL argument");
  *out = nullptr;
  try {
    std::string pqrs;
    for (const M2Mol& m : parse_mol2(lines_of(text, len))) pqrs += pqrs_block(m);
    auto* s = new mmo_molset;
    s->mols = parse_pqrs(pqrs.data(), pqrs.size(), true);
    s->pqrs = std::move(pqrs);
    *out = s;
    return MMO_OK;
  } catch (const InputError& e) {
    return mmo::host_fail(MMO_ERR_BAD_ARG, e.what());
  } catch (const std::bad_alloc&) {
    return mmo::host_fail(MMO_ERR_NOMEM, "out of host memory");
  }
}

int mmo_molset_from_pqrs(const char* text, size_t len, int ligands, mmo_molset** out) {
  if (!out || (!text && len)) return mmo::host_fail(MMO_ERR_BAD_ARG, "NULL argument");
  *out = nullptr;
  try {
    auto* s = new mmo_molset;
    try {
      s->mols = parse_pqrs(text, len, ligands != 0);
    } catch (...) {
      delete s;
      throw;
    }
    s->pqrs.assign(text, len);
    *out = s;
    return MMO_OK;
  } catch (const InputError& e) {
    return mmo::host_fail(MMO_ERR_BAD_ARG, e.what());
  } catch (const std::bad_alloc&) {
    return mmo::host_fail(MMO_ERR_NOMEM, "out of host memory");
  }
}

int mmo_molset_count(const mmo_molset* s) { return s ? (int)s->mols.size() : MMO_ERR_BAD_ARG; }

const char* mmo_molset_name(const mmo_molset* s, int i) {
  if (!s || i < 0 || i >= (int)s->mols.size()) return nullptr;
  return s->mols[i].name.c_str();
}

int mmo_molset_dims(const mmo_molset* s, int i, int* n_atoms, int* n_rbonds, int* n_rg_atoms) {
  if (!s || i < 0 || i >= (int)s->mols.size())
    return mmo::host_fail(MMO_ERR_BAD_ARG, "molecule index out of range");
  const Molecule& m = s->mols[i];
  if (n_atoms) *n_atoms = (int)m.x.size();
  if (n_rbonds) *n_rbonds = m.ligand ? (int)m.rb_left.size() : -1;
  if (n_rg_atoms) *n_rg_atoms = (int)m.rg_atoms.size();
  return MMO_OK;
}

int mmo_molset_atoms(const mmo_molset* s, int i, double* xs, double* ys, double* zs, double* q,
                     double* r, int32_t* anum) {
  if (!s || i < 0 || i >= (int)s->mols.size())
    return mmo::host_fail(MMO_ERR_BAD_ARG, "molecule index out of range");
  const Molecule& m = s->mols[i];
  const size_t n = m.x.size();
  auto put = [n](auto* dst, const auto& src) {
    if (dst && n) std::memcpy(dst, src.data(), n * sizeof(src[0]));
  };
  put(xs, m.x);
  put(ys, m.y);
  put(zs, m.z);
  put(q, m.q);
  put(r, m.r);
  put(anum, m.anum);
  return MMO_OK;
}

int mmo_molset_topology(const mmo_molset* s, int i, int32_t* dists, int32_t* rb_left,
                        int32_t* rb_right, int32_t* rg_off, int32_t* rg_atoms) {
  if (!s || i < 0 || i >= (int)s->mols.size())
    return mmo::host_fail(MMO_ERR_BAD_ARG, "molecule index out of range");
  const Molecule& m = s->mols[i];
  if (!m.ligand) return mmo::host_fail(MMO_ERR_BAD_ARG, "a receptor record has no topology");
  auto put = [](int32_t* dst, const std::vector<int32_t>& src) {
    if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(int32_t));
  };
  put(dists, m.dists);
  put(rb_left, m.rb_left);
  put(rb_right, m.rb_right);
  put(rg_off, m.rg_off);
  put(rg_atoms, m.rg_atoms);
  return MMO_OK;
}

int mmo_molset_pqrs(const mmo_molset* s, char* buf, size_t cap, size_t* len) {
  if (!s || !len) return mmo::host_fail(MMO_ERR_BAD_ARG, "NULL argument");
  *len = s->pqrs.size();
  if (buf) {
    if (cap < s->pqrs.size()) return mmo::host_fail(MMO_ERR_BAD_ARG, "buffer too small");
    std::memcpy(buf, s->pqrs.data(), s->pqrs.size());
  }
  return MMO_OK;
}

void mmo_molset_destroy(mmo_molset* s) { delete s; }

}  // extern "C"
