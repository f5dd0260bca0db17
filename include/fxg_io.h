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
