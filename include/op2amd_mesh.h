/*
 * Synthetic airfoil-channel meshes for the OP2 airfoil workload.
 *
 * The reference reads its mesh from a `new_grid.dat` text file
 * (apps/c/airfoil/airfoil_plain/dp/airfoil.cpp:88-143) written by a MATLAB
 * generator (apps/mesh_generators/naca0012.m:320-345).  That NACA mesh is a
 * network download, so this library generates structured quad meshes with the
 * same conventions instead:
 *   - cell nodes counter-clockwise (naca0012.m:217-228);
 *   - interior edge nodes (n0,n1) oriented so that the normal (dy,-dx) with
 *     (dx,dy) = x[n0]-x[n1] points from ecell[0] to ecell[1]
 *     (naca0012.m:240-259, res_calc.h:68-69);
 *   - boundary edges oriented so that the same normal points out of the domain,
 *     bound = 1 on solid walls, 2 on the far field (naca0012.m:268-303).
 *
 * Domain [0,2]x[0,1], I x J cells, walls at y=0 and y=1, far field at x=0 and
 * x=2.  Interior nodes are jittered uniformly by +-0.2 h per coordinate.
 *
 * Random numbers: SplitMix64 (Steele, Lea, Flood 2014), one stream seeded with
 * `seed`, consumed in this order: jitter (x then y per interior node, j-major),
 * then the cell permutation, then the interior-edge permutation (both
 * Fisher-Yates from the top index down, bounded draws by 64x64->128 multiply
 * high).  Doubles in [0,1) are (r >> 11) * 2^-53.
 *
 * Orderings:
 *   OP2AMD_ORDER_STRUCTURED  j-major base numbering;
 *   OP2AMD_ORDER_RANDOM      cells and interior edges permuted by the seed;
 *   OP2AMD_ORDER_RCM         RANDOM, then reverse Cuthill-McKee on the cell
 *                            graph, propagated to the other sets by the
 *                            reference's rules (op2/src/externlib/op_renumber.cpp:72-151).
 *   OP2AMD_ORDER_HILBERT     RANDOM, then cells sorted by the Hilbert-curve
 *                            index of their structured position (i, j) on the
 *                            smallest 2^k x 2^k grid holding the mesh, the
 *                            other sets propagated as for RCM.
 */
#ifndef OP2AMD_MESH_H
#define OP2AMD_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OP2AMD_ORDER_STRUCTURED = 0,
  OP2AMD_ORDER_RANDOM = 1,
  OP2AMD_ORDER_RCM = 2,
  OP2AMD_ORDER_HILBERT = 3
};

/* All arrays are row-major AoS, 0-based, as the reference's host arrays
 * (airfoil.cpp:99-110). */
typedef struct {
  int nnode, ncell, nedge, nbedge;
  double *x;    /* [nnode][2] */
  int *cell;    /* [ncell][4]  cell -> node */
  int *edge;    /* [nedge][2]  edge -> node */
  int *ecell;   /* [nedge][2]  edge -> cell */
  int *bedge;   /* [nbedge][2] bedge -> node */
  int *becell;  /* [nbedge]    bedge -> cell */
  int *bound;   /* [nbedge]    1 wall, 2 far field */
} op2amd_mesh;

/* Generate an I x J mesh.  Returns 0 on success, nonzero on bad arguments or
 * allocation failure. The arrays are owned by the caller: release them with
 * op2amd_mesh_free. */
int op2amd_mesh_generate(int I, int J, uint64_t seed, int ordering,
                         op2amd_mesh *out);

void op2amd_mesh_free(op2amd_mesh *m);

/* Write / read the reference's text format (airfoil.cpp:94-141): a header
 * "nnode ncell nedge nbedge", then node coordinates, cell->node rows, edge rows
 * "n0 n1 c0 c1", bedge rows "n0 n1 c bound".  Coordinates are written with 17
 * significant digits so that the reference's fscanf("%lf") reads back the
 * exact doubles.  Return 0 on success. */
int op2amd_mesh_write(const op2amd_mesh *m, const char *path);
int op2amd_mesh_read(const char *path, op2amd_mesh *out);

/* Reverse Cuthill-McKee renumbering of an existing mesh (in place), with the
 * propagation rules of op_renumber.cpp:72-151:
 *   cells : RCM on the cell graph (cells adjacent through an interior edge);
 *           start vertex from a George-Liu pseudo-peripheral search that
 *           begins at the lowest-degree, lowest-index cell; neighbours are
 *           queued by (degree, index); the Cuthill-McKee order is reversed.
 *           Disconnected components are ordered one after another.
 *   edges : stable sort by the new index of ecell[e][0] (op_renumber.cpp:84-92)
 *   bedges: stable sort by the new index of becell[b]
 *   nodes : first visit through cell->node in new cell order
 *           (op_renumber.cpp:101-124).
 * The inverse permutations (new index of each old element) are written to
 * the optional output arrays when they are non-NULL. Returns 0 on success. */
int op2amd_mesh_renumber_rcm(op2amd_mesh *m, int *cell_perm, int *edge_perm,
                             int *bedge_perm, int *node_perm);

/* Bandwidth of the cell graph: max over interior edges of |c0 - c1|, and the
 * mean of |c0 - c1| (op_renumber.cpp:369-379 reports the same statistic). */
void op2amd_mesh_cell_bandwidth(const op2amd_mesh *m, int *max_bw,
                                double *mean_bw);

#ifdef __cplusplus
}
#endif

#endif
