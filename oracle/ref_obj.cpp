/*
 * ref_obj.cpp — TEST INFRASTRUCTURE ONLY.
 *
 * Drives the reference's own OBJ loader (src/examples/ObjData.cpp, compiled where it lies by
 * oracle/Makefile; nothing is copied): ObjData::loadFromFile(path).toVertexArray(vdata,
 * idata), with the vertex array written as 8 floats per vertex {vertex, normal, texcoord}.
 * Pins softwarerenderer_amd/obj.py and include/swr/ObjData.h.
 */
#include "ObjData.h"

#include <cstddef>
#include <cstdint>
#include <vector>

extern "C" int swr_ref_obj_load(const char *path, float *vout, size_t vcap, int32_t *iout,
                                size_t icap, size_t *nv, size_t *ni)
{
    std::vector<ObjData::VertexArrayData> vd;
    std::vector<int> id;
    ObjData::loadFromFile(path).toVertexArray(vd, id);
    *nv = vd.size();
    *ni = id.size();
    if (vd.size() > vcap || id.size() > icap)
        return -5;
    for (size_t k = 0; k < vd.size(); ++k) {
        const ObjData::VertexArrayData &v = vd[k];
        const float row[8] = {v.vertex.x, v.vertex.y, v.vertex.z, v.normal.x,
                              v.normal.y, v.normal.z, v.texcoord.x, v.texcoord.y};
        for (int c = 0; c < 8; ++c)
            vout[8 * k + c] = row[c];
    }
    for (size_t k = 0; k < id.size(); ++k)
        iout[k] = id[k];
    return 0;
}
