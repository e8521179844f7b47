#include "mesh.hpp"

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace nsfem
{
    namespace
    {
        struct GmshNode
        {
            long id;
            double x, y;
        };

        struct GmshElement
        {
            long id;
            int entity;
            std::vector<long> points;
        };

        struct GmshName
        {
            int id;
            std::string name;
        };

        // Splits the file into "$Name ... $EndName" blocks (gmsh.cpp:41-95).
        std::map<std::string, std::string> splitSections(std::istream & in)
        {
            std::map<std::string, std::string> sections;
            std::string line, current;
            std::ostringstream body;
            bool inside = false;
            while (std::getline(in, line))
            {
                if (!line.empty() && line.back() == '\r')
                {
                    line.pop_back();
                }
                if (line.empty())
                {
                    continue;
                }
                if (line[0] == '$')
                {
                    if (line.rfind("$End", 0) == 0)
                    {
                        if (inside)
                        {
                            sections[current] += body.str();
                        }
                        inside = false;
                    }
                    else
                    {
                        current = line.substr(1);
                        body.str("");
                        body.clear();
                        inside = true;
                    }
                    continue;
                }
                if (inside)
                {
                    body << line << '\n';
                }
            }
            return sections;
        }

        std::istringstream nextLine(std::istringstream & in, const char * what)
        {
            std::string line;
            if (!std::getline(in, line))
            {
                throw std::runtime_error(std::string("gmsh: unexpected end of section while reading ") + what);
            }
            return std::istringstream(line);
        }

        std::vector<GmshNode> parseNodes(const std::string & text)
        {
            std::istringstream in(text);
            long numBlocks = 0, numNodes = 0, minId = 0, maxId = 0;
            {
                auto h = nextLine(in, "Nodes header");
                if (!(h >> numBlocks >> numNodes >> minId >> maxId))
                {
                    throw std::invalid_argument("Failed to parse header for Nodes");
                }
            }
            std::vector<GmshNode> nodes;
            nodes.reserve(numNodes);
            for (long b = 0; b < numBlocks; b++)
            {
                int dim = 0, tag = 0, parametric = 0;
                long count = 0;
                auto h = nextLine(in, "Nodes block header");
                if (!(h >> dim >> tag >> parametric >> count))
                {
                    throw std::invalid_argument("Failed to parse node block header");
                }
                const size_t first = nodes.size();
                for (long j = 0; j < count; j++)
                {
                    auto l = nextLine(in, "node tag");
                    GmshNode n{};
                    l >> n.id;
                    nodes.push_back(n);
                }
                for (long j = 0; j < count; j++)
                {
                    auto l = nextLine(in, "node coordinates");
                    double x = 0, y = 0, z = 0;
                    l >> x >> y >> z;
                    // el::Point stores float: round the input the same way.
                    nodes[first + j].x = static_cast<double>(static_cast<float>(x));
                    nodes[first + j].y = static_cast<double>(static_cast<float>(y));
                }
            }
            return nodes;
        }

        std::vector<GmshElement> parseElements(const std::string & text)
        {
            std::istringstream in(text);
            long numBlocks = 0, numElements = 0, minId = 0, maxId = 0;
            {
                auto h = nextLine(in, "Elements header");
                if (!(h >> numBlocks >> numElements >> minId >> maxId))
                {
                    throw std::invalid_argument("Failed to parse header for Elements");
                }
            }
            std::vector<GmshElement> elements;
            elements.reserve(numElements);
            for (long b = 0; b < numBlocks; b++)
            {
                int dim = 0, tag = 0, type = 0;
                long count = 0;
                auto h = nextLine(in, "Elements block header");
                if (!(h >> dim >> tag >> type >> count))
                {
                    throw std::invalid_argument("Failed to parse element block header");
                }
                for (long j = 0; j < count; j++)
                {
                    auto l = nextLine(in, "element");
                    GmshElement e{};
                    e.entity = tag;
                    if (!(l >> e.id))
                    {
                        throw std::invalid_argument("Failed to parse element tag");
                    }
                    long p;
                    while (l >> p)
                    {
                        e.points.push_back(p);
                    }
                    elements.push_back(std::move(e));
                }
            }
            return elements;
        }

        // Curve entity tag -> first physical tag (points/surfaces are skipped,
        // as the reference only maps curves: Mesh/src/io.cpp:106-112).
        std::map<int, int> parseCurvePhysicalTags(const std::string & text)
        {
            std::istringstream in(text);
            long nPoints = 0, nCurves = 0, nSurfaces = 0, nVolumes = 0;
            {
                auto h = nextLine(in, "Entities header");
                if (!(h >> nPoints >> nCurves >> nSurfaces >> nVolumes))
                {
                    throw std::runtime_error("Failed to parse entities header");
                }
            }
            for (long i = 0; i < nPoints; i++)
            {
                nextLine(in, "point entity");
            }
            std::map<int, int> result;
            for (long i = 0; i < nCurves; i++)
            {
                auto l = nextLine(in, "curve entity");
                int tag = 0;
                double bb[6];
                long nPhys = 0;
                if (!(l >> tag >> bb[0] >> bb[1] >> bb[2] >> bb[3] >> bb[4] >> bb[5] >> nPhys))
                {
                    throw std::runtime_error("Failed to parse curve");
                }
                if (nPhys > 0)
                {
                    int phys = 0;
                    l >> phys;
                    result[tag] = phys;
                }
            }
            return result;
        }

        std::vector<GmshName> parsePhysicalNames(const std::string & text)
        {
            std::istringstream in(text);
            long n = 0;
            {
                auto h = nextLine(in, "PhysicalNames header");
                if (!(h >> n))
                {
                    throw std::invalid_argument("Failed to parse number of names in PhysicalSection");
                }
            }
            std::vector<GmshName> names;
            for (long i = 0; i < n; i++)
            {
                auto l = nextLine(in, "physical name");
                int dim = 0, tag = 0;
                std::string name;
                if (!(l >> dim >> tag >> name))
                {
                    throw std::runtime_error("Failed to parse line from physics section");
                }
                if (name.size() < 2)
                {
                    throw std::runtime_error("Name too short");
                }
                names.push_back({tag, name.substr(1, name.size() - 2)});
            }
            return names;
        }

        // Attaches oriented border edges to (triangle, side) pairs by scanning
        // triangles in order and trying all 6 oriented sides (io.cpp:161-205).
        void attachBorders(TriangleMesh & mesh, const std::vector<std::vector<std::array<int, 3>>> & edges)
        {
            // edges[from] = list of {to, group, claimed}
            std::vector<std::vector<std::array<int, 3>>> e = edges;
            const int nElements = static_cast<int>(mesh.elements.size());
            for (int i = 0; i < nElements; i++)
            {
                const auto & pts = mesh.elements[i];
                for (int side = 0; side < 6; side++)
                {
                    const int from = pts[kTriangleSides[side].from];
                    const int to = pts[kTriangleSides[side].to];
                    for (auto & edge : e[from])
                    {
                        if (edge[0] == to)
                        {
                            if (edge[2])
                            {
                                throw std::runtime_error("Border edge claimed twice");
                            }
                            edge[2] = 1;
                            mesh.borderElements.push_back(BorderElement{i, side, edge[1]});
                            break;
                        }
                    }
                }
            }
        }
    } // namespace

    TriangleMesh parseTriangleGmsh(const std::string & fileName)
    {
        std::ifstream file(fileName);
        if (!file.is_open())
        {
            throw std::runtime_error("Failed to open mesh file [" + fileName + "]");
        }
        auto sections = splitSections(file);
        auto need = [&](const char * name) -> const std::string &
        {
            auto it = sections.find(name);
            if (it == sections.end())
            {
                throw std::runtime_error(std::string("No ") + name + " section!");
            }
            return it->second;
        };

        auto nodes = parseNodes(need("Nodes"));
        auto elements = parseElements(need("Elements"));
        auto curveTags = parseCurvePhysicalTags(need("Entities"));
        auto names = parsePhysicalNames(need("PhysicalNames"));

        std::sort(nodes.begin(), nodes.end(), [](const auto & a, const auto & b) { return a.id < b.id; });
        std::sort(elements.begin(), elements.end(), [](const auto & a, const auto & b) { return a.id < b.id; });
        std::sort(names.begin(), names.end(), [](const auto & a, const auto & b) { return a.id < b.id; });

        auto contiguous = [](const auto & v)
        {
            for (size_t i = 1; i < v.size(); i++)
            {
                if (v[i].id != v[i - 1].id + 1)
                {
                    return false;
                }
            }
            return true;
        };
        if (!contiguous(nodes))
        {
            throw std::runtime_error("Node IDs are not contiguous");
        }
        if (!contiguous(elements))
        {
            throw std::runtime_error("Element IDs are not contiguous");
        }
        if (!contiguous(names))
        {
            throw std::runtime_error("Physical group IDs are not contiguous");
        }
        if (nodes.empty() || names.empty())
        {
            throw std::runtime_error("Mesh has no nodes or no physical groups");
        }

        const long nodeOffset = nodes[0].id;
        const int groupOffset = names[0].id;

        TriangleMesh mesh;
        mesh.nodes.resize(nodes.size());
        for (size_t i = 0; i < nodes.size(); i++)
        {
            mesh.nodes[i] = Point{nodes[i].x, nodes[i].y};
        }
        mesh.groups.resize(names.size());
        for (const auto & n : names)
        {
            mesh.groups[n.id - groupOffset] = n.name;
        }

        const int numNodes = static_cast<int>(nodes.size());
        std::vector<std::vector<std::array<int, 3>>> edges(numNodes);
        size_t parsed = 0;
        for (const auto & el : elements)
        {
            if (el.points.size() != 2)
            {
                continue;
            }
            parsed++;
            const int from = static_cast<int>(el.points[0] - nodeOffset);
            const int to = static_cast<int>(el.points[1] - nodeOffset);
            int group = 0; // std::map::operator[] default in the reference
            auto it = curveTags.find(el.entity);
            if (it != curveTags.end())
            {
                group = it->second - groupOffset;
            }
            if (from < 0 || from >= numNodes || to < 0 || to >= numNodes)
            {
                throw std::runtime_error("Border element refers to a missing node");
            }
            edges[from].push_back({to, group, 0});
        }
        for (const auto & el : elements)
        {
            if (el.points.size() != 3)
            {
                continue;
            }
            parsed++;
            std::array<int, 3> tri;
            for (int k = 0; k < 3; k++)
            {
                tri[k] = static_cast<int>(el.points[k] - nodeOffset);
                if (tri[k] < 0 || tri[k] >= numNodes)
                {
                    throw std::runtime_error("Triangle refers to a missing node");
                }
            }
            mesh.elements.push_back(tri);
        }
        if (parsed != elements.size())
        {
            throw std::runtime_error("Not all elements have been parsed");
        }

        attachBorders(mesh, edges);
        return mesh;
    }

    TriangleMesh buildTriangleMesh(const std::vector<Point> & nodes,
                                   const std::vector<std::array<int, 3>> & triangles,
                                   const std::vector<std::array<int, 3>> & borderEdges,
                                   const std::vector<std::string> & groups)
    {
        TriangleMesh mesh;
        mesh.nodes = nodes;
        mesh.elements = triangles;
        mesh.groups = groups;
        const int numNodes = static_cast<int>(nodes.size());
        std::vector<std::vector<std::array<int, 3>>> edges(numNodes);
        for (const auto & e : borderEdges)
        {
            if (e[0] < 0 || e[0] >= numNodes || e[1] < 0 || e[1] >= numNodes)
            {
                throw std::invalid_argument("Border edge refers to a missing node");
            }
            if (e[2] < 0 || e[2] >= static_cast<int>(groups.size()))
            {
                throw std::invalid_argument("Border edge has an invalid group");
            }
            edges[e[0]].push_back({e[1], e[2], 0});
        }
        for (const auto & t : triangles)
        {
            for (int k = 0; k < 3; k++)
            {
                if (t[k] < 0 || t[k] >= numNodes)
                {
                    throw std::invalid_argument("Triangle refers to a missing node");
                }
            }
        }
        attachBorders(mesh, edges);
        return mesh;
    }

    int ConcreteMesh::findGroupId(const std::string & name) const
    {
        auto it = std::find(groups.begin(), groups.end(), name);
        return it == groups.end() ? -1 : static_cast<int>(it - groups.begin());
    }

    ConcreteMesh createMesh(const TriangleMesh & tri, ElementType type)
    {
        ConcreteMesh result;
        result.type = type;
        result.ptsPerSide = type == ElementType::P2 ? 3 : 2;
        result.nodesPerElement = type == ElementType::P2 ? 6 : 3;
        result.nodes = tri.nodes;
        const int nVertices = static_cast<int>(tri.nodes.size());

        // Edge graph keyed by the smaller vertex, sides kept in first-seen order
        // (concreteMesh.cpp:153-192).
        struct Side
        {
            int to;
            int extra;
        };
        std::vector<std::vector<Side>> graph(nVertices);
        for (const auto & el : tri.elements)
        {
            for (int a = 0; a < 3; a++)
            {
                const int pa = el[a];
                const int pb = el[(a + 1) % 3];
                if (pa == pb)
                {
                    throw std::invalid_argument("Degenerate triangle side");
                }
                const int from = std::min(pa, pb);
                const int to = std::max(pa, pb);
                bool present = false;
                for (const auto & s : graph[from])
                {
                    if (s.to == to)
                    {
                        present = true;
                        break;
                    }
                }
                if (!present)
                {
                    graph[from].push_back(Side{to, -1});
                }
            }
        }

        // One extra (midpoint) node per side for P2, numbered by ascending
        // smaller vertex then insertion order (concreteMesh.cpp:194-217).
        if (type == ElementType::P2)
        {
            for (int from = 0; from < nVertices; from++)
            {
                const Point a = result.nodes[from];
                for (auto & s : graph[from])
                {
                    const Point b = result.nodes[s.to];
                    s.extra = static_cast<int>(result.nodes.size());
                    result.nodes.push_back(Point{0.5 * a.x + 0.5 * b.x, 0.5 * a.y + 0.5 * b.y});
                }
            }
        }

        auto sideExtra = [&](int a, int b) -> int
        {
            const int from = std::min(a, b);
            const int to = std::max(a, b);
            for (const auto & s : graph[from])
            {
                if (s.to == to)
                {
                    return s.extra;
                }
            }
            throw std::runtime_error("createMesh: missing side");
        };

        result.numElements = static_cast<int>(tri.elements.size());
        const int npe = result.nodesPerElement;
        result.elements.resize(static_cast<size_t>(npe) * result.numElements);
        result.elementTransforms.resize(result.numElements);
        const int sideStep = result.ptsPerSide - 1;
        for (int i = 0; i < result.numElements; i++)
        {
            const auto & src = tri.elements[i];
            const Point p0 = tri.nodes[src[0]];
            const Point p1 = tri.nodes[src[1]];
            const Point p2 = tri.nodes[src[2]];
            AffineMap & t = result.elementTransforms[i];
            t.b[0] = p0.x;
            t.b[1] = p0.y;
            t.m[0][0] = p1.x - p0.x;
            t.m[0][1] = p2.x - p0.x;
            t.m[1][0] = p1.y - p0.y;
            t.m[1][1] = p2.y - p0.y;

            int * ids = result.elements.data() + static_cast<size_t>(i) * npe;
            for (int c = 0; c < 3; c++)
            {
                ids[c * sideStep] = src[c];
            }
            if (type == ElementType::P2)
            {
                for (int a = 0; a < 3; a++)
                {
                    ids[a * sideStep + 1] = sideExtra(src[a], src[(a + 1) % 3]);
                }
            }
        }

        // Border elements: triangle, side, group, then the side's nodes from
        // the side's start vertex to its end vertex (concreteMesh.cpp:323-350).
        result.numBorderElements = static_cast<int>(tri.borderElements.size());
        const int stride = result.borderElementStride();
        result.borderElements.resize(static_cast<size_t>(stride) * result.numBorderElements);
        for (int i = 0; i < result.numBorderElements; i++)
        {
            const auto & be = tri.borderElements[i];
            const auto & el = tri.elements[be.element];
            const int idA = el[kTriangleSides[be.side].from];
            const int idB = el[kTriangleSides[be.side].to];
            int * dst = result.borderElements.data() + static_cast<size_t>(i) * stride;
            dst[0] = be.element;
            dst[1] = be.side;
            dst[2] = be.group;
            dst[3] = idA;
            dst[stride - 1] = idB;
            if (type == ElementType::P2)
            {
                dst[4] = sideExtra(idA, idB);
            }
        }
        result.groups = tri.groups;
        return result;
    }
} // namespace nsfem
