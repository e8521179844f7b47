// Finite elements, quadrature, local matrices and CSR assembly (host, FP64).
//
// Follows the reference's Element and Fem libraries:
//   P1 / P2 shape functions       cpp/libs/Element/src/p1.cpp:35-57, p2.cpp:35-84
//   degree-4 Cowper rule          cpp/libs/Element/src/triangleIntegrator.cpp:27-35
//   local mass / stiffness        triangleIntegrator.cpp:132-190
//   local self-convection         triangleIntegrator.cpp:192-231
//   local divergence              triangleIntegrator.cpp:233-293
//   CSR prototypes + assembly     cpp/libs/Fem/src/chorinCsr.cpp:142-502
//   FastConvection map            cpp/libs/Fem/src/fastConvection.cpp:9-115
// All arithmetic here is FP64.  The reference evaluates the same integrals in
// FP32; patterns are identical, values agree to FP32 rounding.
#pragma once

#include <array>
#include <vector>

#include "csr.hpp"
#include "mesh.hpp"

namespace nsfem
{
    struct QuadPoint
    {
        double x, y, w;
    };

    // Degree-4, 6-point rule on the reference triangle (0,0),(1,0),(0,1).
    const std::vector<QuadPoint> & quadratureDeg4();

    void p1Shape(double x, double y, double * phi);
    void p1Grad(double x, double y, double * gx, double * gy);
    void p2Shape(double x, double y, double * phi);
    void p2Grad(double x, double y, double * gx, double * gy);

    // B = [[m11, -m10], [-m01, m00]] so that grad_x = B grad_ref / det.
    struct ElementGeometry
    {
        double b00, b01, b10, b11;
        double det;
    };
    ElementGeometry geometryOf(const AffineMap & t);

    // Local matrices, row-major.  nDof = 3 (P1) or 6 (P2).
    void localMass(const AffineMap & t, int nDof, double * dst);
    void localStiffness(const AffineMap & t, int nDof, double * dst);
    // P2 convection with the P2 flow (ux, uy) given at the 6 element nodes.
    void localSelfConvectionP2(const AffineMap & t, const double * ux, const double * uy, double * dst);
    // swap = false: rows = P1 (3), cols = P2 (6):  X[r][c] = int d(phi2_c)/dx psi1_r
    // swap = true : rows = P2 (6), cols = P1 (3):  X[r][c] = int d(psi1_c)/dx phi2_r
    void localDivergence(const AffineMap & t, bool swap, double * dstX, double * dstY);

    // Reference-element convection tensors: Tx[r][c][k] = sum_q w phi_r phi_k dphi_c/dxi,
    // Ty likewise with d/deta (P2, 6x6x6 each).  For an element with geometry g:
    //   C[r][c] = sgn(det) * sum_k (g.b00*ux_k + g.b10*uy_k) Tx[r][c][k] + (g.b01*ux_k + g.b11*uy_k) Ty[r][c][k]
    struct ConvectionTensors
    {
        double tx[6][6][6];
        double ty[6][6][6];
    };
    const ConvectionTensors & convectionTensors();

    struct ChorinMatrices
    {
        CsrD velocityMass;        // N x N
        CsrD velocityStiffness;   // N x N
        CsrD pressureStiffness;   // V x V
        CsrD velocityPressureDiv; // V x 2N
        CsrD pressureVelocityDiv; // 2N x V
    };

    ChorinMatrices buildChorinMatrices(const ConcreteMesh & velocityMesh, const ConcreteMesh & pressureMesh, int nThreads);

    // The FastConvection integration map: rows = entries of the velocity CSR
    // pattern, cols = 2N; C.values = Int * [ux; uy].
    CsrD buildFastConvectionMap(const ConcreteMesh & velocityMesh, const CsrD & velocityPattern, int nThreads);

    // Element-wise convection data for the GPU kernel.
    struct ConvectionElementData
    {
        int numElements = 0;
        std::vector<int> elementNodes;    // E * 6 (P2 ids)
        std::vector<double> geometry;     // E * 5: b00, b01, b10, b11, sgn(det)
        std::vector<int> contribStart;    // nnz + 1
        std::vector<int> contribIndex;    // 36 E: e*36 + r*6 + c, grouped by CSR entry
    };
    // slots (optional): the velocity slot lists of buildAssemblyTopology,
    // which are exactly the contributor lists (not recomputed then).
    struct SlotList;
    ConvectionElementData buildConvectionElementData(const ConcreteMesh & velocityMesh, const CsrD & velocityPattern,
                                                     const SlotList * slots = nullptr);

    // Topology of the Chorin matrices for assembly elsewhere (the GPU): the
    // CSR patterns (values empty) and, for every entry, the element-local
    // slots that land on it, in ascending element order (the order the host
    // assembly sums them in).  Slots: velocity e*36 + a*6 + b (P2 local a, b);
    // pressure e*9 + a*3 + b; VPD e*36 + comp*18 + a*6 + b (P1 row a, P2 col
    // b, comp 0 = x, 1 = y); PVD e*36 + comp*18 + a*3 + b (P2 row a, P1 col b).
    struct SlotList
    {
        std::vector<int> start; // nnz + 1
        std::vector<int> slot;
    };
    struct AssemblyTopology
    {
        CsrD velocity, pressure, velocityPressureDiv, pressureVelocityDiv; // patterns
        SlotList velocitySlots, pressureSlots, vpdSlots, pvdSlots;
    };
    AssemblyTopology buildAssemblyTopology(const ConcreteMesh & velocityMesh, const ConcreteMesh & pressureMesh, int nThreads);

    int defaultThreads();
} // namespace nsfem
