// swr/ObjData.h — Wavefront OBJ input for the draw path, with the interface of the
// reference examples' ObjData (ObjData.h:31-57, ObjData.cpp:69-200): loadFromFile() then
// toVertexArray() gives a vertex array {vertex, normal, texcoord} (32 bytes per vertex:
// bind it with stride sizeof(VertexArrayData) and the SWR_VS_MVP_POS_UV program) and
// triangle indices.  Header-only, host code.
//
// Kept from the reference: 1-based position / normal / texcoord lists with an implicit
// zero entry 0 (a missing "/t" or "/n" picks the zero vector); one output vertex per
// distinct (position, normal, texcoord) triple, numbered in order of first appearance;
// faces fanned around their first corner; unknown line types ignored; numbers read with
// std::istream >> (a missing component is 0; the reference leaves a missing texture
// coordinate component uninitialised).  Where the reference reads out of bounds (an
// index past its list, a face with no corners, more than two slashes in a corner),
// toVertexArray / loadFromFile throw std::runtime_error.  The reference's ObjData is a
// global class; this one is swr::ObjData (`using swr::ObjData;` restores the spelling).
#pragma once

#include <cstdint>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

namespace swr {

struct ObjVec3 {
    float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct ObjVec2 {
    float x = 0.0f, y = 0.0f;
};

struct ObjData {
    struct VertexArrayData {
        ObjVec3 vertex;
        ObjVec3 normal;
        ObjVec2 texcoord;
    };

    struct VertexRef {
        unsigned vertexIndex = 0;
        unsigned normalIndex = 0;
        unsigned texcoordIndex = 0;
    };

    typedef std::vector<VertexRef> Face;

    std::vector<ObjVec3> vertices{ObjVec3()};
    std::vector<ObjVec3> normals{ObjVec3()};
    std::vector<ObjVec2> texcoords{ObjVec2()};
    std::vector<Face> faces;

    static ObjData loadFromFile(const char *filename)
    {
        std::ifstream in(filename);
        if (!in)
            throw std::runtime_error(std::string("cannot open ") + filename);
        return parse(in);
    }

    static ObjData parse(std::istream &in)
    {
        ObjData r;
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            std::string cmd;
            if (!(ls >> cmd) || cmd[0] == '#')
                continue;
            if (cmd == "v" || cmd == "vn") {
                ObjVec3 v;
                ls >> v.x >> v.y >> v.z;
                (cmd == "v" ? r.vertices : r.normals).push_back(v);
            } else if (cmd == "vt") {
                ObjVec2 t;
                ls >> t.x >> t.y;
                r.texcoords.push_back(t);
            } else if (cmd == "f") {
                Face face;
                std::string word;
                while (ls >> word)
                    face.push_back(corner(word));
                if (face.empty())
                    throw std::runtime_error("OBJ face with no corners");
                r.faces.push_back(face);
            }
        }
        return r;
    }

    void toVertexArray(std::vector<VertexArrayData> &vdata, std::vector<int> &idata) const
    {
        vdata.clear();
        idata.clear();
        std::map<std::tuple<unsigned, unsigned, unsigned>, int> index_of;
        auto add = [&](const VertexRef &c) {
            const auto key = std::make_tuple(c.vertexIndex, c.normalIndex, c.texcoordIndex);
            const auto it = index_of.find(key);
            if (it != index_of.end())
                return it->second;
            if (c.vertexIndex >= vertices.size() || c.normalIndex >= normals.size() ||
                c.texcoordIndex >= texcoords.size())
                throw std::runtime_error("OBJ corner indexes past its list");
            const int k = (int)index_of.size();
            index_of.emplace(key, k);
            vdata.push_back({vertices[c.vertexIndex], normals[c.normalIndex],
                             texcoords[c.texcoordIndex]});
            return k;
        };
        for (const Face &f : faces) {
            const int i1 = add(f[0]);
            for (size_t j = 2; j < f.size(); ++j) {
                const int i2 = add(f[j - 1]);
                const int i3 = add(f[j]);
                idata.push_back(i1);
                idata.push_back(i2);
                idata.push_back(i3);
            }
        }
    }

private:
    // "v/t/n" (any part may be empty): vertex, normal and texcoord index, 0 when absent
    static VertexRef corner(const std::string &w)
    {
        std::string part[3];
        int k = 0;
        for (char ch : w) {
            if (ch == '/') {
                if (++k > 2)
                    throw std::runtime_error("OBJ face corner with more than two slashes");
            } else {
                part[k] += ch;
            }
        }
        VertexRef r;
        r.vertexIndex = index(part[0]);
        r.texcoordIndex = index(part[1]);
        r.normalIndex = index(part[2]);
        return r;
    }

    static unsigned index(const std::string &s)
    {
        std::istringstream is(s);
        long long v = 0;
        if (!(is >> v))
            return 0;
        if (v < 0 || v > 0xFFFFFFFFll)
            throw std::runtime_error("OBJ index out of range");
        return (unsigned)v;
    }
};

static_assert(sizeof(ObjData::VertexArrayData) == 32, "VertexArrayData is 8 floats");

}  // namespace swr
