// mesh_shim.cpp -- C entry point over the reference's own mesh library
// (jukan_koodit/Mesh), compiled in place into oracle/_ref/libmeshref.so by
// oracle/Makefile.  TEST INFRASTRUCTURE ONLY: tests/ use it to check the
// 14-point primal operator of the oracle against the DEC Laplacian the
// reference builds on a BCC grid.
//
// For every node of DelaunayMesh::createBccGrid(minp, maxp, h)
// (DelaunayMesh.cpp:178) it reports the node position and, per incident
// edge, the offset to the other node and the weight
//   getEdgeHodge(e) / getNodeHodge(n)          (Mesh.cpp:1479-1489),
// i.e. the row of star0^-1 d0^T star1 d0 for that node.
#include "Mesh/DelaunayMesh.hpp"

extern "C" int ref_bcc_stencils(double h, double ext, int max_nodes, int max_nb, double *pos, int *count,
                                double *nb) {
    DelaunayMesh m(3);
    m.createBccGrid(Vector3(-ext, -ext, -ext), Vector3(ext, ext, ext), h);
    const int n = (int)m.getNodeSize();
    if (n > max_nodes) return -n;
    for (int i = 0; i < n; i++) {
        const Vector4 p = m.getNodePosition(i);
        pos[3 * i] = p.x;
        pos[3 * i + 1] = p.y;
        pos[3 * i + 2] = p.z;
        const Buffer<uint> &e = m.getNodeEdges(i);
        const double nh = m.getNodeHodge(i);
        int c = 0;
        for (uint q = 0; q < e.size() && c < max_nb; q++, c++) {
            const Vector4 d = m.getNodePosition(m.getEdgeOtherNode(e[q], i)) - p;
            double *o = nb + 4 * ((long long)i * max_nb + c);
            o[0] = d.x;
            o[1] = d.y;
            o[2] = d.z;
            o[3] = m.getEdgeHodge(e[q]) / nh;
        }
        count[i] = (int)e.size();
    }
    return n;
}
