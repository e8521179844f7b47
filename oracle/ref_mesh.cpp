// ORACLE - test infrastructure only.
//
// A C entry point over the reference's own P1/P2 numbering, mesh::createMesh
// (cpp/libs/Mesh/src/concreteMesh.cpp:78-293) with the el::P1 / el::P2
// elements (cpp/libs/Element/src/p1.cpp, p2.cpp, element.cpp, point.cpp,
// affineTransform.cpp).  oracle/Makefile compiles those reference sources
// where they lie, together with this file, into oracle/_ref/libref_mesh.so;
// tests compare the oracle's and the library's numbering with it.
#include <element/p1.h>
#include <element/p2.h>
#include <mesh/concreteMesh.h>

#include <cstring>
#include <exception>

extern "C" {

// Builds the reference ConcreteMesh of order 1 (P1) or 2 (P2) from a
// triangle mesh: xy (2 * numNodes floats), triangles (3 * numTriangles),
// border (3 * numBorder: triangle, side, group; the TriangleMesh border
// elements).  Outputs: outXy (2 * nodeCapacity floats), outElements
// (numTriangles * nodes per element), outBorder (numBorder * (3 + points per
// side)), *outNumNodes.  Returns 0, or -1 if nodeCapacity is too small, -2 on
// an exception.
int ref_create_mesh(int order, int numNodes, const float * xy, int numTriangles, const int * triangles, int numBorder,
                    const int * border, int numGroups, int nodeCapacity, float * outXy, int * outElements, int * outBorder,
                    int * outNumNodes)
{
    try
    {
        mesh::TriangleMesh tm;
        tm.nodes.resize(numNodes);
        for (int i = 0; i < numNodes; i++)
        {
            tm.nodes[i] = el::Point{xy[2 * i], xy[2 * i + 1]};
        }
        tm.elements.resize(numTriangles);
        for (int e = 0; e < numTriangles; e++)
        {
            tm.elements[e] = {triangles[3 * e], triangles[3 * e + 1], triangles[3 * e + 2]};
        }
        for (int g = 0; g < numGroups; g++)
        {
            tm.groups.push_back("group" + std::to_string(g));
        }
        tm.borderElements.resize(numBorder);
        for (int b = 0; b < numBorder; b++)
        {
            tm.borderElements[b] = mesh::BorderElement{border[3 * b], border[3 * b + 1], border[3 * b + 2]};
        }
        static const el::P1 p1;
        static const el::P2 p2;
        const el::Element & base = order == 1 ? static_cast<const el::Element &>(p1) : static_cast<const el::Element &>(p2);
        const mesh::ConcreteMesh cm = mesh::createMesh(tm, base);
        const int n = static_cast<int>(cm.nodes.size());
        *outNumNodes = n;
        if (n > nodeCapacity)
        {
            return -1;
        }
        for (int i = 0; i < n; i++)
        {
            outXy[2 * i] = cm.nodes[i].x;
            outXy[2 * i + 1] = cm.nodes[i].y;
        }
        std::memcpy(outElements, cm.elements.data(), cm.elements.size() * sizeof(int));
        std::memcpy(outBorder, cm.borderElements.data(), cm.borderElements.size() * sizeof(int));
        return 0;
    }
    catch (const std::exception &)
    {
        return -2;
    }
}
}
