/* fxg_io: FLEXI's mesh and state files (HDF5) for the fxg library, host side only.
 *
 * Reads HOPR mesh files the way FLEXI's ReadMesh / setLocalSideIDs / fillMeshInfo do
 * (src/mesh/mesh_readin.f90:197-560, src/mesh/prepare_mesh.f90:50-737) and returns the arrays
 * fxg_create / fxg_create_from_nodes take: SideToElem(5,nSides), ElemToSide(2,6,nElems) and
 * NodeCoords(3,0:NGeo,0:NGeo,0:NGeo,nElems), all Fortran-ordered and 1-based like FLEXI's. Only
 * conforming hexahedral meshes whose sides are all inner or periodic are accepted (the sides the
 * fxg kernels handle); anything else returns an error.
 *
 * Writes and reads FLEXI State files: dataset DG_Solution(5,0:N,0:N,0:N,nElems) and the header
 * attributes of GenerateFileSkeleton / WriteHeader (src/io_hdf5/hdf5_output.f90:741-847,
 * 940-960), so FLEXI can restart from a state written here (src/restart/restart.f90:120-441)
 * and this library can read FLEXI's.
 *
 * Every function returns 0 on success; fxg_io_last_error() describes a failure.
 * Built against the HDF5 C library (libfxg_io.so); the GPU library libfxg.so does not need it.
 */
#ifndef FXG_IO_H
#define FXG_IO_H

#ifdef __cplusplus
extern "C" {
#endif

const char* fxg_io_last_error(void);

/* Sizes of a HOPR mesh file: attributes nElems and Ngeo, and the number of distinct sides
 * (every side of a conforming periodic mesh is counted once). Replaces the size queries of
 * ReadMesh (mesh_readin.f90:262-309). */
int fxg_io_mesh_size(const char* path, int* nElems, int* nSides, int* NGeo);

/* Connectivity and geometry of a HOPR mesh file (ReadMesh + setLocalSideIDs + fillMeshInfo):
 * SideIDs numbered in element/local-side order of first appearance (prepare_mesh.f90:199-243),
 * the side's flip-0 element is its master (prepare_mesh.f90:705-716). Arrays are caller-owned,
 * sized from fxg_io_mesh_size. */
int fxg_io_read_mesh(const char* path, int nElems, int nSides, int NGeo, int* SideToElem, int* ElemToSide,
                     double* NodeCoords);

/* The same for meshes with boundary sides (ReadBCs + ReadMesh + setLocalSideIDs, mesh_readin.f90:
 * 70-184, 397-470, prepare_mesh.f90:196-243). userNames[nUser] / userTypes[2*nUser] are the
 * BoundaryName / BoundaryType (type, state) pairs of FLEXI's parameter file, which replace the type
 * and state of the file's boundary of the same name (case-insensitive); nUser may be 0. Sides of
 * boundaries whose type is not 0, 1 or 100 are boundary sides: numbered first (1..nBCSides), flip 0,
 * no neighbour (SideToElem(2,:) = -1). nBCs is the number of boundaries in the file. */
int fxg_io_mesh_size_bc(const char* path, int nUser, const char* const* userNames, const int* userTypes, int* nElems,
                        int* nSides, int* NGeo, int* nBCSides, int* nBCs);
/* Arrays as fxg_io_read_mesh plus BC(nBCSides) (the boundary index of each boundary side, fillMeshInfo
 * prepare_mesh.f90:717) and BoundaryType(nBCs,3) = (type, state, alpha) in FLEXI's Fortran layout
 * (mesh_readin.f90:168-173); BCNames may be NULL, else nBCs strings of nameLen bytes each. */
int fxg_io_read_mesh_bc(const char* path, int nUser, const char* const* userNames, const int* userTypes, int nElems,
                        int nSides, int NGeo, int nBCSides, int nBCs, int* SideToElem, int* ElemToSide,
                        double* NodeCoords, int* BC, int* BoundaryType, char* BCNames, int nameLen);

/* Sizes of a State file's DG_Solution (GetDataProps, hdf5_input.f90:307-360). */
int fxg_io_state_size(const char* path, int* nVar, int* N, int* nElems);

/* DG_Solution(nVar,0:N,0:N,0:N,nElems) into U, the attribute Time into *time and the attributes
 * NodeType and MeshFile (trailing blanks removed) into the given buffers (restart.f90:255-441). */
int fxg_io_read_state(const char* path, int nVar, int N, int nElems, double* U, double* time, char* nodeType,
                      int nodeTypeLen, char* meshFile, int meshFileLen);

/* A FLEXI State file (WriteState, hdf5_output.f90:60-188): DG_Solution(5,0:N,0:N,0:N,nElems)
 * from U and the attributes Program, File_Type "State", Project_Name, File_Version, N,
 * NComputation, Dimension 3, Time, MeshFile, NextFile, NodeType, VarNames. */
int fxg_io_write_state(const char* path, const char* projectName, const char* meshFile, const char* nodeType, int N,
                       int nElems, double time, const double* U);

#ifdef __cplusplus
}
#endif
#endif
