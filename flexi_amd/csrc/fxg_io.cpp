// FLEXI mesh (HOPR) and State files through the HDF5 C API: include/fxg_io.h.
//
// Host-only code (built with g++ into libfxg_io.so). The mesh reader follows ReadMesh
// (src/mesh/mesh_readin.f90:197-470), setLocalSideIDs (src/mesh/prepare_mesh.f90:50-243) and
// fillMeshInfo (src/mesh/prepare_mesh.f90:679-720) for the conforming, single-partition case.
#include "../../include/fxg_io.h"

#include <hdf5.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& msg) {
  g_err = msg;
  return 1;
}

// RAII handles of the HDF5 objects opened here.
struct H {
  hid_t id;
  herr_t (*close)(hid_t);
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0) close(id);
  }
  H(const H&) = delete;
  H& operator=(const H&) = delete;
};

// Silence HDF5's own error printing; failures are reported through fxg_io_last_error.
struct Quiet {
  H5E_auto2_t f = nullptr;
  void* d = nullptr;
  Quiet() {
    H5Eget_auto2(H5E_DEFAULT, &f, &d);
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  }
  ~Quiet() { H5Eset_auto2(H5E_DEFAULT, f, d); }
};

int read_int_attr(hid_t file, const char* name, int* v) {
  H a(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
  if (a.id < 0) return fail(std::string("attribute ") + name + " missing");
  if (H5Aread(a.id, H5T_NATIVE_INT, v) < 0) return fail(std::string("cannot read attribute ") + name);
  return 0;
}

int read_double_attr(hid_t file, const char* name, double* v) {
  H a(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
  if (a.id < 0) return fail(std::string("attribute ") + name + " missing");
  if (H5Aread(a.id, H5T_NATIVE_DOUBLE, v) < 0) return fail(std::string("cannot read attribute ") + name);
  return 0;
}

// First element of a string attribute (fixed-length, blank- or null-padded), trailing blanks off.
int read_str_attr(hid_t file, const char* name, std::string* out) {
  H a(H5Aopen(file, name, H5P_DEFAULT), H5Aclose);
  if (a.id < 0) return fail(std::string("attribute ") + name + " missing");
  H t(H5Aget_type(a.id), H5Tclose);
  if (H5Tget_class(t.id) != H5T_STRING || H5Tis_variable_str(t.id) > 0)
    return fail(std::string("attribute ") + name + " is not a fixed-length string");
  H s(H5Aget_space(a.id), H5Sclose);
  const hssize_t n = H5Sget_simple_extent_npoints(s.id);
  const size_t len = H5Tget_size(t.id);
  std::vector<char> buf((size_t)(n > 0 ? n : 1) * len + 1, '\0');
  if (H5Aread(a.id, t.id, buf.data()) < 0) return fail(std::string("cannot read attribute ") + name);
  std::string v(buf.data(), len);
  const size_t z = v.find('\0');
  if (z != std::string::npos) v.resize(z);
  while (!v.empty() && v.back() == ' ') v.pop_back();
  *out = v;
  return 0;
}

// A 2D dataset of int (C dims rows x cols) into out.
int read_int_2d(hid_t file, const char* name, size_t cols, std::vector<int>& out, size_t* rows) {
  H d(H5Dopen2(file, name, H5P_DEFAULT), H5Dclose);
  if (d.id < 0) return fail(std::string("dataset ") + name + " missing");
  H s(H5Dget_space(d.id), H5Sclose);
  hsize_t dims[2] = {0, 0};
  if (H5Sget_simple_extent_ndims(s.id) != 2 || H5Sget_simple_extent_dims(s.id, dims, nullptr) != 2 || dims[1] != cols)
    return fail(std::string("dataset ") + name + " has an unexpected shape");
  out.resize((size_t)dims[0] * cols);
  *rows = (size_t)dims[0];
  if (H5Dread(d.id, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) < 0)
    return fail(std::string("cannot read dataset ") + name);
  return 0;
}

// One element's local side as ReadMesh sets it up (mesh_readin.f90:312-392, 397-460).
struct LocSide {
  int ind = 0;       // |SideInfo(SIDE_ID)|
  int flip = 0;      // 0 for the oriented (master) side
  int nbElem = 0;    // SideInfo(SIDE_nbElemID)
  int bc = 0;        // SideInfo(SIDE_BCID)
  bool boundary = false;  // a boundary side without a connection (BC type other than 0, 1, 100)
  int sideID = -1;   // assigned by setLocalSideIDs
  int conn = -1;     // index (6*elem + locSide) of the connected side
};

struct MeshData {
  int nElems = 0, NGeo = 0, nSides = 0, nBCSides = 0;
  std::vector<LocSide> sides;  // 6 per element
  std::vector<double> nodes;   // (3,0:NGeo,0:NGeo,0:NGeo,nElems)
  std::vector<std::string> bcNames;
  std::vector<int> bcType;     // BCType(4,nBCs) after the parameter file's BoundaryType
};

// BoundaryName / BoundaryType pairs of the parameter file (ReadBCs, mesh_readin.f90:92-100).
struct UserBCs {
  int n = 0;
  const char* const* names = nullptr;
  const int* types = nullptr;  // (type, state) per name
};

std::string lower_trim(std::string v) {
  while (!v.empty() && (v.back() == ' ' || v.back() == '\0')) v.pop_back();
  for (char& ch : v) ch = (char)std::tolower((unsigned char)ch);
  return v;
}

// BCNames(nBCs) (fixed-length strings, blank padded).
int read_bc_names(hid_t file, size_t nBCs, std::vector<std::string>& out) {
  H d(H5Dopen2(file, "BCNames", H5P_DEFAULT), H5Dclose);
  if (d.id < 0) return fail("dataset BCNames missing");
  H t(H5Dget_type(d.id), H5Tclose);
  if (H5Tget_class(t.id) != H5T_STRING || H5Tis_variable_str(t.id) > 0)
    return fail("BCNames is not a fixed-length string dataset");
  H s(H5Dget_space(d.id), H5Sclose);
  if (H5Sget_simple_extent_npoints(s.id) != (hssize_t)nBCs) return fail("BCNames and BCType differ in length");
  const size_t len = H5Tget_size(t.id);
  std::vector<char> buf(len * nBCs + 1, '\0');
  if (H5Dread(d.id, t.id, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()) < 0) return fail("cannot read BCNames");
  out.clear();
  for (size_t i = 0; i < nBCs; ++i) {
    std::string v(&buf[i * len], len);
    const size_t z = v.find('\0');
    if (z != std::string::npos) v.resize(z);
    while (!v.empty() && v.back() == ' ') v.pop_back();
    out.push_back(v);
  }
  return 0;
}

// ReadBCs (mesh_readin.f90:113-164): a user boundary replaces the type and state of the file's
// boundary of the same name (case-insensitive); every user name must exist; a non-periodic
// boundary cannot become periodic.
int apply_user_bcs(const UserBCs& u, const std::vector<std::string>& names, std::vector<int>& bct) {
  std::vector<int> map(names.size(), -1);
  for (int k = 0; k < u.n; ++k) {
    if (!u.names || !u.names[k] || !u.types) return fail("NULL user boundary name or type");
    bool found = false;
    for (size_t i = 0; i < names.size(); ++i)
      if (lower_trim(names[i]) == lower_trim(u.names[k])) {
        if (map[i] >= 0) return fail("boundary " + names[i] + " is redefined multiple times");
        map[i] = k;
        found = true;
      }
    if (!found) return fail(std::string("boundary condition not found in the mesh file: ") + u.names[k]);
  }
  for (size_t i = 0; i < names.size(); ++i) {
    if (map[i] < 0) continue;
    const int type = u.types[2 * map[i]], state = u.types[2 * map[i] + 1];
    if (type == 1 && bct[4 * i] != 1) return fail("remapping non-periodic to periodic boundaries is not possible");
    bct[4 * i] = type;
    bct[4 * i + 2] = state;
  }
  return 0;
}

int load_mesh(const char* path, MeshData& md, bool withNodes, bool allowBC = false, const UserBCs& user = UserBCs{}) {
  Quiet q;
  H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return fail(std::string("cannot open mesh file ") + path);
  int K = 0, NGeo = 0;
  if (read_int_attr(f.id, "nElems", &K) || read_int_attr(f.id, "Ngeo", &NGeo)) return 1;
  if (K < 1 || NGeo < 1) return fail("mesh file: nElems and Ngeo must be positive");
  int isMortar = 0;
  if (H5Aexists(f.id, "isMortarMesh") > 0 && read_int_attr(f.id, "isMortarMesh", &isMortar) == 0 && isMortar == 1)
    return fail("mortar meshes are not supported (conforming meshes only)");
  // boundary conditions: BCType(4,nBCs) = (type, curve index, state, alpha) (readBCs)
  std::vector<int> bct;
  size_t nBCs = 0;
  if (H5Lexists(f.id, "BCType", H5P_DEFAULT) > 0 && read_int_2d(f.id, "BCType", 4, bct, &nBCs)) return 1;
  md.bcNames.clear();
  if (nBCs > 0 && read_bc_names(f.id, nBCs, md.bcNames)) return 1;
  if (user.n > 0 && apply_user_bcs(user, md.bcNames, bct)) return 1;
  md.bcType = bct;
  std::vector<int> ei, si;
  size_t rowsE = 0, rowsS = 0;
  if (read_int_2d(f.id, "ElemInfo", 6, ei, &rowsE) || read_int_2d(f.id, "SideInfo", 5, si, &rowsS)) return 1;
  if ((int)rowsE != K) return fail("mesh file: ElemInfo rows differ from nElems");
  md.nElems = K;
  md.NGeo = NGeo;
  md.sides.assign((size_t)6 * K, LocSide{});
  for (int e = 0; e < K; ++e) {
    const int* E = &ei[(size_t)6 * e];
    if (E[0] % 10 != 8)  // HOPR hexahedra: 108, 118, 208
      return fail("only hexahedral elements are supported (ElemType " + std::to_string(E[0]) + ")");
    if (E[3] - E[2] != 6) return fail("element with other than six sides (mortar or non-hexahedral)");
    if (E[5] - E[4] != (NGeo + 1) * (NGeo + 1) * (NGeo + 1))
      return fail("element node count differs from (Ngeo+1)^3");
    for (int ls = 0; ls < 6; ++ls) {
      const size_t r = (size_t)E[2] + ls;
      if (r >= rowsS) return fail("SideInfo index out of range");
      const int* S = &si[5 * r];
      if (S[0] < 0 || S[2] < 0) return fail("mortar sides are not supported (conforming meshes only)");
      LocSide& a = md.sides[(size_t)6 * e + ls];
      a.ind = S[1] > 0 ? S[1] : -S[1];
      a.flip = S[1] > 0 ? 0 : S[3] % 10;
      if (a.flip < 0 || a.flip > 4) return fail("side flip out of range");
      a.nbElem = S[2];
      a.bc = S[4];
      if (a.bc != 0) {
        if (a.bc < 0 || (size_t)a.bc > nBCs) return fail("SideInfo BCID out of range");
        const int type = bct[(size_t)4 * (a.bc - 1)];
        // sides of other types are boundary sides: no connection, flip 0 (mesh_readin.f90:413-422)
        if (type != 0 && type != 1 && type != 100) {
          if (!allowBC)
            return fail("boundary sides are not supported (BC type " + std::to_string(type) +
                        "): only inner and periodic sides");
          a.boundary = true;
          a.flip = 0;
          continue;
        }
      }
      if (a.nbElem == 0) return fail("side without a neighbour element and without a boundary condition");
      if (a.nbElem > K) return fail("neighbour element index out of range");
    }
  }
  // connections: the neighbour's side with the same index (mesh_readin.f90:433-448)
  for (int e = 0; e < K; ++e)
    for (int ls = 0; ls < 6; ++ls) {
      LocSide& a = md.sides[(size_t)6 * e + ls];
      if (a.conn >= 0 || a.boundary) continue;
      const int nb = a.nbElem - 1;
      for (int ms = 0; ms < 6; ++ms) {
        const int j = 6 * nb + ms;
        if (j == 6 * e + ls) continue;
        if (md.sides[j].ind == a.ind) {
          a.conn = j;
          md.sides[j].conn = 6 * e + ls;
          break;
        }
      }
      if (a.conn < 0) return fail("side " + std::to_string(a.ind) + " has no partner in its neighbour element");
    }
  // side IDs in element / local-side order (prepare_mesh.f90:196-243): boundary sides first
  // (1..nBCSides), then the inner sides
  int nBC = 0;
  for (const LocSide& a : md.sides) nBC += a.boundary ? 1 : 0;
  int nextBC = 0, next = nBC;
  for (int e = 0; e < K; ++e)
    for (int ls = 0; ls < 6; ++ls) {
      LocSide& a = md.sides[(size_t)6 * e + ls];
      if (a.sideID >= 0) continue;
      if (a.boundary) {
        a.sideID = ++nextBC;
        continue;
      }
      LocSide& b = md.sides[a.conn];
      if ((a.flip == 0) == (b.flip == 0)) return fail("side " + std::to_string(a.ind) + " needs one master and one slave");
      a.sideID = b.sideID = ++next;
    }
  md.nSides = next;
  md.nBCSides = nBC;
  if (withNodes) {
    H d(H5Dopen2(f.id, "NodeCoords", H5P_DEFAULT), H5Dclose);
    if (d.id < 0) return fail("dataset NodeCoords missing");
    H s(H5Dget_space(d.id), H5Sclose);
    hsize_t dims[2] = {0, 0};
    const hsize_t nn = (hsize_t)K * (NGeo + 1) * (NGeo + 1) * (NGeo + 1);
    if (H5Sget_simple_extent_ndims(s.id) != 2 || H5Sget_simple_extent_dims(s.id, dims, nullptr) != 2 || dims[1] != 3 ||
        dims[0] < nn)
      return fail("dataset NodeCoords has an unexpected shape");
    // each element's nodes are contiguous from ElemInfo(FirstNodeInd) (mesh_readin.f90:471-490)
    std::vector<double> all((size_t)dims[0] * 3);
    if (H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, all.data()) < 0)
      return fail("cannot read dataset NodeCoords");
    const size_t per = (size_t)3 * (NGeo + 1) * (NGeo + 1) * (NGeo + 1);
    md.nodes.resize(per * K);
    for (int e = 0; e < K; ++e) {
      // FirstNodeInd is an offset read from the file: reject negative values, and compare
      // without forming 3*first + per, which could wrap around
      const long long f = ei[(size_t)6 * e + 4];
      if (f < 0 || (unsigned long long)f > all.size() / 3 || per > all.size() - 3 * (size_t)f)
        return fail("NodeCoords index out of range");
      const size_t first = (size_t)f;
      std::memcpy(&md.nodes[per * e], &all[3 * first], per * sizeof(double));
    }
  }
  return 0;
}

// fillMeshInfo (prepare_mesh.f90:690-716): SideToElem initialised to -1 as in InitMesh
// (mesh.f90:295-297), the flip-0 element of a side is its master, BC(SideID) for boundary sides.
void fill_mesh_info(const MeshData& md, int* SideToElem, int* ElemToSide, double* NodeCoords, int* BC) {
  for (size_t i = 0; i < (size_t)5 * md.nSides; ++i) SideToElem[i] = -1;
  for (int e = 0; e < md.nElems; ++e)
    for (int ls = 0; ls < 6; ++ls) {
      const LocSide& a = md.sides[(size_t)6 * e + ls];
      ElemToSide[2 * ls + 12 * (size_t)e] = a.sideID;
      ElemToSide[2 * ls + 1 + 12 * (size_t)e] = a.flip;
      int* s2e = SideToElem + (size_t)5 * (a.sideID - 1);
      if (a.flip == 0) {
        s2e[0] = e + 1;
        s2e[2] = ls + 1;
      } else {
        s2e[1] = e + 1;
        s2e[3] = ls + 1;
        s2e[4] = a.flip;
      }
      if (BC && a.boundary) BC[a.sideID - 1] = a.bc;
    }
  std::memcpy(NodeCoords, md.nodes.data(), md.nodes.size() * sizeof(double));
}

hid_t str_type(size_t len) {
  const hid_t t = H5Tcopy(H5T_C_S1);
  H5Tset_size(t, len > 0 ? len : 1);
  H5Tset_strpad(t, H5T_STR_SPACEPAD);
  return t;
}

int write_str_attr(hid_t obj, const char* name, const std::vector<std::string>& vals, size_t len) {
  H t(str_type(len), H5Tclose);
  const hsize_t n = vals.size();
  H s(H5Screate_simple(1, &n, nullptr), H5Sclose);
  std::vector<char> buf(len * vals.size(), ' ');
  for (size_t i = 0; i < vals.size(); ++i) std::memcpy(&buf[i * len], vals[i].data(), std::min(len, vals[i].size()));
  H a(H5Acreate2(obj, name, t.id, s.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (a.id < 0 || H5Awrite(a.id, t.id, buf.data()) < 0) return fail(std::string("cannot write attribute ") + name);
  return 0;
}

template <class T>
int write_num_attr(hid_t obj, const char* name, hid_t memType, hid_t fileType, const T* v, hsize_t n) {
  H s(H5Screate_simple(1, &n, nullptr), H5Sclose);
  H a(H5Acreate2(obj, name, fileType, s.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose);
  if (a.id < 0 || H5Awrite(a.id, memType, v) < 0) return fail(std::string("cannot write attribute ") + name);
  return 0;
}

std::string base_name(const std::string& p) {
  const size_t k = p.find_last_of('/');
  return k == std::string::npos ? p : p.substr(k + 1);
}

}  // namespace

extern "C" {

const char* fxg_io_last_error(void) { return g_err.c_str(); }

int fxg_io_mesh_size(const char* path, int* nElems, int* nSides, int* NGeo) {
  if (!path || !nElems || !nSides || !NGeo) return fail("NULL argument");
  MeshData md;
  if (load_mesh(path, md, false)) return 1;
  *nElems = md.nElems;
  *nSides = md.nSides;
  *NGeo = md.NGeo;
  return 0;
}

int fxg_io_read_mesh(const char* path, int nElems, int nSides, int NGeo, int* SideToElem, int* ElemToSide,
                     double* NodeCoords) {
  if (!path || !SideToElem || !ElemToSide || !NodeCoords) return fail("NULL argument");
  MeshData md;
  if (load_mesh(path, md, true)) return 1;
  if (md.nElems != nElems || md.nSides != nSides || md.NGeo != NGeo)
    return fail("array sizes differ from the mesh file (call fxg_io_mesh_size first)");
  fill_mesh_info(md, SideToElem, ElemToSide, NodeCoords, nullptr);
  return 0;
}

int fxg_io_mesh_size_bc(const char* path, int nUser, const char* const* userNames, const int* userTypes, int* nElems,
                        int* nSides, int* NGeo, int* nBCSides, int* nBCs) {
  if (!path || !nElems || !nSides || !NGeo || !nBCSides || !nBCs || nUser < 0) return fail("invalid argument");
  MeshData md;
  if (load_mesh(path, md, false, true, UserBCs{nUser, userNames, userTypes})) return 1;
  *nElems = md.nElems;
  *nSides = md.nSides;
  *NGeo = md.NGeo;
  *nBCSides = md.nBCSides;
  *nBCs = (int)md.bcNames.size();
  return 0;
}

int fxg_io_read_mesh_bc(const char* path, int nUser, const char* const* userNames, const int* userTypes, int nElems,
                        int nSides, int NGeo, int nBCSides, int nBCs, int* SideToElem, int* ElemToSide,
                        double* NodeCoords, int* BC, int* BoundaryType, char* BCNames, int nameLen) {
  if (!path || !SideToElem || !ElemToSide || !NodeCoords || nUser < 0) return fail("invalid argument");
  if ((nBCSides > 0 && !BC) || (nBCs > 0 && !BoundaryType)) return fail("BC or BoundaryType is NULL");
  MeshData md;
  if (load_mesh(path, md, true, true, UserBCs{nUser, userNames, userTypes})) return 1;
  if (md.nElems != nElems || md.nSides != nSides || md.NGeo != NGeo || md.nBCSides != nBCSides ||
      (int)md.bcNames.size() != nBCs)
    return fail("array sizes differ from the mesh file (call fxg_io_mesh_size_bc first)");
  fill_mesh_info(md, SideToElem, ElemToSide, NodeCoords, BC);
  // BoundaryType(nBCs,3) = (type, state, alpha), mesh_readin.f90:168-173
  for (int i = 0; i < nBCs; ++i) {
    BoundaryType[i] = md.bcType[(size_t)4 * i];
    BoundaryType[i + nBCs] = md.bcType[(size_t)4 * i + 2];
    BoundaryType[i + 2 * nBCs] = md.bcType[(size_t)4 * i + 3];
    if (BCNames && nameLen > 0) std::snprintf(BCNames + (size_t)i * nameLen, (size_t)nameLen, "%s", md.bcNames[i].c_str());
  }
  return 0;
}

int fxg_io_state_size(const char* path, int* nVar, int* N, int* nElems) {
  if (!path || !nVar || !N || !nElems) return fail("NULL argument");
  Quiet q;
  H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return fail(std::string("cannot open state file ") + path);
  H d(H5Dopen2(f.id, "DG_Solution", H5P_DEFAULT), H5Dclose);
  if (d.id < 0) return fail("dataset DG_Solution missing");
  H s(H5Dget_space(d.id), H5Sclose);
  hsize_t dims[5] = {0, 0, 0, 0, 0};
  if (H5Sget_simple_extent_ndims(s.id) != 5 || H5Sget_simple_extent_dims(s.id, dims, nullptr) != 5)
    return fail("DG_Solution is not a 5-dimensional array (3D states only)");
  if (dims[1] != dims[2] || dims[1] != dims[3]) return fail("DG_Solution is not (nVar,0:N,0:N,0:N,nElems)");
  *nElems = (int)dims[0];
  *N = (int)dims[1] - 1;
  *nVar = (int)dims[4];
  int Na = 0;
  if (read_int_attr(f.id, "N", &Na)) return 1;
  if (Na != *N) return fail("attribute N differs from the DG_Solution size");
  return 0;
}

int fxg_io_read_state(const char* path, int nVar, int N, int nElems, double* U, double* time, char* nodeType,
                      int nodeTypeLen, char* meshFile, int meshFileLen) {
  if (!path || !U || !time) return fail("NULL argument");
  int v = 0, n = 0, k = 0;
  if (fxg_io_state_size(path, &v, &n, &k)) return 1;
  if (v != nVar || n != N || k != nElems) return fail("array sizes differ from the state file");
  Quiet q;
  H f(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return fail(std::string("cannot open state file ") + path);
  H d(H5Dopen2(f.id, "DG_Solution", H5P_DEFAULT), H5Dclose);
  if (d.id < 0 || H5Dread(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, U) < 0)
    return fail("cannot read DG_Solution");
  if (read_double_attr(f.id, "Time", time)) return 1;
  std::string s;
  if (nodeType && nodeTypeLen > 0) {
    if (read_str_attr(f.id, "NodeType", &s)) return 1;
    std::snprintf(nodeType, (size_t)nodeTypeLen, "%s", s.c_str());
  }
  if (meshFile && meshFileLen > 0) {
    if (read_str_attr(f.id, "MeshFile", &s)) return 1;
    std::snprintf(meshFile, (size_t)meshFileLen, "%s", s.c_str());
  }
  return 0;
}

int fxg_io_write_state(const char* path, const char* projectName, const char* meshFile, const char* nodeType, int N,
                       int nElems, double time, const double* U) {
  if (!path || !projectName || !meshFile || !nodeType || !U) return fail("NULL argument");
  if (N < 1 || nElems < 1) return fail("N and nElems must be positive");
  Quiet q;
  H f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (f.id < 0) return fail(std::string("cannot create state file ") + path);
  // header (WriteHeader) and the state attributes (GenerateFileSkeleton)
  const double version = 0.1;  // FileVersion, src/output/output_vars.f90:30
  const int dim = 3;
  const std::string proj(projectName), mesh(meshFile);
  if (write_str_attr(f.id, "Program", {"Flexi"}, 5) || write_str_attr(f.id, "File_Type", {"State"}, 5) ||
      write_str_attr(f.id, "Project_Name", {proj}, proj.size()) ||
      write_num_attr(f.id, "File_Version", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, &version, 1) ||
      write_num_attr(f.id, "N", H5T_NATIVE_INT, H5T_STD_I32LE, &N, 1) ||
      write_num_attr(f.id, "Dimension", H5T_NATIVE_INT, H5T_STD_I32LE, &dim, 1) ||
      write_num_attr(f.id, "Time", H5T_NATIVE_DOUBLE, H5T_IEEE_F64LE, &time, 1) ||
      write_str_attr(f.id, "MeshFile", {mesh}, mesh.size()) ||
      write_str_attr(f.id, "NextFile", {base_name(path)}, 255) ||
      write_str_attr(f.id, "NodeType", {std::string(nodeType)}, 255) ||
      write_num_attr(f.id, "NComputation", H5T_NATIVE_INT, H5T_STD_I32LE, &N, 1) ||
      write_str_attr(f.id, "VarNames", {"Density", "MomentumX", "MomentumY", "MomentumZ", "EnergyStagnationDensity"},
                     255))
    return 1;
  // DG_Solution(5,0:N,0:N,0:N,nElems): Fortran order, so the HDF5 (C order) dims are reversed
  const hsize_t dims[5] = {(hsize_t)nElems, (hsize_t)N + 1, (hsize_t)N + 1, (hsize_t)N + 1, 5};
  H s(H5Screate_simple(5, dims, nullptr), H5Sclose);
  H d(H5Dcreate2(f.id, "DG_Solution", H5T_IEEE_F64LE, s.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Dclose);
  if (d.id < 0 || H5Dwrite(d.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, U) < 0)
    return fail("cannot write DG_Solution");
  return 0;
}

}  // extern "C"
