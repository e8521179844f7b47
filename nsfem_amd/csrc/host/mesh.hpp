// Triangle meshes and their P1/P2 node numbering (host side, set-up only).
//
// Follows the reference's Mesh library:
//   gmsh 4.1 ASCII ingest      cpp/libs/Mesh/src/gmsh.cpp:41-499
//   TriangleMesh + borders     cpp/libs/Mesh/src/io.cpp:65-208, include/mesh/triMesh.h
//   P1/P2 numbering            cpp/libs/Mesh/src/concreteMesh.cpp:78-293
// Node numbering decides every CSR pattern and the graph colouring, so the
// numbering rules are kept exactly: vertices keep gmsh order, edge nodes are
// created per (min vertex, insertion order of first sight) and border elements
// are discovered by scanning triangles x 6 oriented sides.
#pragma once

#include <array>
#include <string>
#include <vector>

namespace nsfem
{
    struct Point
    {
        double x, y;
    };

    struct TriangleSide
    {
        int from, to;
    };

    // Oriented triangle sides, indexed by BorderElement::side (triMesh.h:16-24).
    inline constexpr std::array<TriangleSide, 6> kTriangleSides = {
        TriangleSide{0, 1}, TriangleSide{1, 0}, TriangleSide{1, 2},
        TriangleSide{2, 1}, TriangleSide{2, 0}, TriangleSide{0, 2}};

    struct BorderElement
    {
        int element; // triangle that owns the border edge
        int side;    // index into kTriangleSides
        int group;   // physical group id (0-based)
    };

    // Corner-only mesh, element agnostic.
    struct TriangleMesh
    {
        std::vector<Point> nodes;
        std::vector<std::array<int, 3>> elements;
        std::vector<std::string> groups;
        std::vector<BorderElement> borderElements;
    };

    // Parses a gmsh 4.1 ASCII file (lines and triangles only).  Coordinates are
    // rounded to float on input, as the reference stores el::Point as float.
    TriangleMesh parseTriangleGmsh(const std::string & fileName);

    // Builds a TriangleMesh from raw arrays: border edges are oriented
    // (from, to) node pairs with a group id; they are attached to triangles the
    // same way the gmsh path attaches line elements.
    TriangleMesh buildTriangleMesh(const std::vector<Point> & nodes,
                                   const std::vector<std::array<int, 3>> & triangles,
                                   const std::vector<std::array<int, 3>> & borderEdges,
                                   const std::vector<std::string> & groups);

    enum class ElementType : int
    {
        P1 = 1,
        P2 = 2
    };

    struct AffineMap
    {
        // x = m * xi + b, m[0][0] = p1.x - p0.x, m[0][1] = p2.x - p0.x, ...
        double m[2][2];
        double b[2];
    };

    struct ConcreteMesh
    {
        ElementType type;
        int nodesPerElement; // 3 (P1) or 6 (P2)
        int ptsPerSide;      // 2 (P1) or 3 (P2)
        std::vector<Point> nodes;
        int numElements = 0;
        std::vector<int> elements; // numElements * nodesPerElement
        std::vector<AffineMap> elementTransforms;
        std::vector<std::string> groups;
        int numBorderElements = 0;
        // numBorderElements * (ptsPerSide + 3): triangle, side, group, pt0, ..., ptLast
        std::vector<int> borderElements;

        int borderElementStride() const { return ptsPerSide + 3; }
        int findGroupId(const std::string & name) const;
    };

    ConcreteMesh createMesh(const TriangleMesh & tri, ElementType type);
} // namespace nsfem
