// Host-only driver of lal_linalg::mtx (include/lal_linalg.hpp) for the CPU
// parity tests against LinAlgLib's own read_mtx / write_mtx:
//
//   mtx_tool read  <file.mtx> <out.bin>   parse; write m, n, nnz, row_ptr,
//                                          col_ind, val as raw int32/float64
//                                          (exit 3 when the file is refused)
//   mtx_tool write <in.bin> <file.mtx>     the inverse: arrays -> text file
//
// No device call is made, so it runs without a GPU.
#include "lal_linalg.hpp"

#include <cstdio>
#include <fstream>

namespace
{
    template <typename T>
    void put(std::ofstream& o, const std::vector<T>& v)
    {
        o.write(reinterpret_cast<const char*>(v.data()), sizeof(T) * v.size());
    }
    template <typename T>
    bool get(std::ifstream& in, std::vector<T>& v, size_t n)
    {
        v.resize(n);
        return bool(in.read(reinterpret_cast<char*>(v.data()), sizeof(T) * n));
    }
}

int main(int argc, char** argv)
{
    if(argc != 4)
    {
        std::fprintf(stderr, "usage: mtx_tool read|write <in> <out>\n");
        return 2;
    }
    const std::string mode = argv[1];
    if(mode == "read")
    {
        int                 m = 0, n = 0;
        std::vector<int>    rp, ci;
        std::vector<double> v;
        if(!lal_linalg::mtx::read(argv[2], m, n, rp, ci, v))
            return 3;
        std::ofstream o(argv[3], std::ios::binary);
        const int     head[3] = {m, n, (int)ci.size()};
        o.write(reinterpret_cast<const char*>(head), sizeof(head));
        put(o, rp);
        put(o, ci);
        put(o, v);
        return o ? 0 : 1;
    }
    if(mode == "write")
    {
        std::ifstream in(argv[2], std::ios::binary);
        int           head[3];
        if(!in.read(reinterpret_cast<char*>(head), sizeof(head)))
            return 1;
        std::vector<int>    rp, ci;
        std::vector<double> v;
        if(!get(in, rp, head[0] + 1) || !get(in, ci, head[2]) || !get(in, v, head[2]))
            return 1;
        return lal_linalg::mtx::write(argv[3], head[0], head[1], rp, ci, v) ? 0 : 1;
    }
    return 2;
}
