/*
 * cfd_mesh.h - host-side mesh generation and median-dual preprocessing for the
 * unstructured finite-volume Euler hot path.
 *
 * This library produces exactly the arrays the reference solver builds before
 * its first residual evaluation:
 *   - UGRID mesh content (nodes, boundary triangles with tags, tetrahedra),
 *     Euler3D/src/MeshIO.cpp:50-218 (reader), :22-44 (winding translation);
 *   - node->node neighbour lists in first-appearance order,
 *     Euler3D/src/Connectivity_Maps.cpp:44-340 and :915-990 (CRS form);
 *   - interior edges (L < R, node-major order), Connectivity_Maps.cpp:733-762;
 *   - boundary edges pairing a physical node with a ghost node nNode+k,
 *     Euler3D/src/Area_Volume.cpp:462-657;
 *   - median-dual edge area vectors and control volumes,
 *     Euler3D/src/Area_Volume.cpp:660-821, outward boundary normals :903-917.
 * The arithmetic of the metrics follows the reference operation by operation,
 * so the arrays are bit-identical to the reference's for the same UGRID input.
 *
 * Everything is O(N + E); the reference's own node->edge build and
 * Cuthill-McKee reorder are O(N^2) and are not used.
 *
 * All functions return 0 on success and a negative CFDM_E_* code on failure;
 * cfdm_last_error() returns a description of the last failure in this thread.
 */
#ifndef CFD_MESH_H
#define CFD_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFDM_OK 0
#define CFDM_E_ARG -1
#define CFDM_E_IO -2
#define CFDM_E_OOM -3
#define CFDM_E_MESH -4

/* Boundary tags written by cfdm_box_generate (UGRID surface tags, 1-based). */
#define CFDM_TAG_XMIN 1
#define CFDM_TAG_XMAX 2
#define CFDM_TAG_YMIN 3 /* lower wall, carries the circular-arc bump */
#define CFDM_TAG_YMAX 4
#define CFDM_TAG_ZMIN 5
#define CFDM_TAG_ZMAX 6

typedef struct cfdm_mesh cfdm_mesh;

/* Counts of a mesh.  The edge/volume fields are valid after cfdm_preprocess. */
typedef struct cfdm_info {
    int32_t n_node;   /* physical nodes */
    int32_t n_tri;    /* boundary triangles */
    int32_t n_tet;    /* tetrahedra */
    int32_t n_edge;   /* interior edges (node pairs) */
    int32_t n_bedge;  /* boundary edges == ghost nodes (3 per triangle) */
    int32_t n_n2n;    /* entries in the node->node CRS */
} cfdm_info;

/*
 * Hex grid [0,lx]x[0,ly]x[0,lz] with nx*ny*nz hexes, each split into 6 Kuhn
 * tetrahedra along its main diagonal.  bump_h > 0 maps the lower wall over
 * x in [1,2] to a circular arc of height bump_h (radius (0.25+h^2)/(2h)),
 * with y = y_b(x) + (j/ny)*(ly - y_b(x)).  jitter > 0 moves every interior
 * node by U(-jitter, +jitter) * spacing per axis, drawn from SplitMix64 keyed
 * on (seed, lexicographic node index).  reorder = 1 renumbers nodes in
 * reverse Cuthill-McKee order (O(N+E)).
 */
int cfdm_box_generate(int nx, int ny, int nz, double lx, double ly, double lz,
                      double bump_h, double jitter, uint64_t seed, int reorder,
                      cfdm_mesh **out);

int cfdm_read_ugrid(const char *path, cfdm_mesh **out);
int cfdm_write_ugrid(const cfdm_mesh *m, const char *path);
void cfdm_free(cfdm_mesh *m);

/* Builds connectivity, edges, metrics (the reference's preprocessing). */
int cfdm_preprocess(cfdm_mesh *m);

int cfdm_get_info(const cfdm_mesh *m, cfdm_info *info);

/*
 * Copies out arrays (any pointer may be NULL).  Layouts match the reference:
 *   xyz[3N]            node coordinates, interleaved x,y,z
 *   tet[4*n_tet], tri[3*n_tri] (after winding translation and orientation
 *                      fix), tri_tag[n_tri]
 *   edge[2E]           interior edge (L,R), L < R
 *   edge_area[3E]      area vector, oriented L -> R
 *   bedge[2B]          (physical node, ghost node = n_node + k)
 *   bedge_area[3B]     outward boundary area vector
 *   bedge_tag[B]       surface tag of the parent triangle
 *   vol[N]             median-dual control volume
 *   n2n_ia[N+1], n2n_ja[n_n2n]  neighbour lists in the reference order
 */
int cfdm_get_arrays(const cfdm_mesh *m, double *xyz, int32_t *tet, int32_t *tri,
                    int32_t *tri_tag, int32_t *edge, double *edge_area,
                    int32_t *bedge, double *bedge_area, int32_t *bedge_tag,
                    double *vol, int32_t *n2n_ia, int32_t *n2n_ja);

/* Restart_Writer (Euler3D/src/RestartIO.cpp:45-109), conservative variables:
 * int32 iteration, node count, ghost count, variable count (5), then the 5
 * variables of each of the n_node + n_ghost nodes interleaved per node.
 * q is [5][n_node + n_ghost] (as cfd_field_get(CFD_FIELD_Q)). */
int cfdm_restart_write(const char *path, int32_t iteration, int32_t n_node, int32_t n_ghost, const double *q);
/* Restart_Reader (Euler3D/src/RestartIO.cpp:112-179): fails with CFDM_E_ARG if
 * the file's node or ghost count differs from the mesh's. */
int cfdm_restart_read(const char *path, int32_t n_node, int32_t n_ghost, int32_t *iteration, double *q);
/* VTK_Writer (Euler3D/src/MeshIO.cpp:219-399), conservative variables: legacy
 * ASCII unstructured grid of the boundary triangles and tetrahedra with
 * BoundaryIds, density, velocity, pressure, temperature and Mach.  q is
 * [5][ldq] with the nodes first; gamma and r_gas of the non-dimensionalisation. */
int cfdm_vtk_write(const cfdm_mesh *m, const char *path, const double *q, int32_t ldq, double gamma, double r_gas);

const char *cfdm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* CFD_MESH_H */
