// Loads an OBJ file with swr::ObjData (include/swr/ObjData.h) and writes
// {u32 vertex count, u32 index count, vertices (8 floats each), indices} to a file.
#include <swr/ObjData.h>

#include <cstdio>

int main(int argc, char **argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: obj_dump in.obj out.bin\n");
        return 2;
    }
    std::vector<swr::ObjData::VertexArrayData> v;
    std::vector<int> idx;
    swr::ObjData::loadFromFile(argv[1]).toVertexArray(v, idx);
    FILE *f = std::fopen(argv[2], "wb");
    if (!f)
        return 1;
    const uint32_t n[2] = {(uint32_t)v.size(), (uint32_t)idx.size()};
    std::fwrite(n, 4, 2, f);
    if (!v.empty())
        std::fwrite(v.data(), sizeof v[0], v.size(), f);
    if (!idx.empty())
        std::fwrite(idx.data(), 4, idx.size(), f);
    std::fclose(f);
    return 0;
}
