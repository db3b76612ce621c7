// Checkpoint / restart of level states (host code around device buffers).
//
// Reference: AmrLevelAdv::checkPoint / restart (Source/AmrLevelAdv.cpp:
// 95-113), which hand the level's state to AMReX's AmrLevel::checkPoint /
// AmrLevel::restart, and the final checkpoint main.cpp writes
// (Source/main.cpp:97-100).  What a restart needs to continue the run bit
// for bit is the same here: the simulated time, each level's step count and
// dt, its box list, and its state.
//
// Layout of a checkpoint directory <dir>:
//   <dir>/Header          text: format tag, dim, level count, time, refinement
//                         ratio, problem extent; per level its step count, dt,
//                         domain box, cell sizes, ghost width, box list, the
//                         state's length in doubles and an FNV-1a 64-bit
//                         checksum of its bytes
//   <dir>/Level_<l>/State the level's state exactly as the device holds it:
//                         the packed FABs of bx_pack_offsets (component-major,
//                         x fastest, ghost cells included), IEEE
//                         little-endian doubles
// Doubles in the header are printed with 17 significant digits, which reads
// back to the same bits.
#include <sys/stat.h>

#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "bx_level.h"

namespace {

const char *const CHK_TAG = "BAMReX-HIP-Checkpoint-1";

// states move through the host in chunks of this many doubles
constexpr size_t CHUNK = (size_t)8 << 20; // 64 MiB

bool make_dir(const std::string &p)
{
    return mkdir(p.c_str(), 0755) == 0 || errno == EEXIST;
}

uint64_t fnv1a(uint64_t h, const void *data, size_t n)
{
    const unsigned char *p = (const unsigned char *)data;
    for (size_t i = 0; i < n; ++i)
    {
        h ^= p[i];
        h *= 1099511628211ULL;
    }
    return h;
}
constexpr uint64_t FNV_OFFSET = 14695981039346656037ULL;

std::string fmt(double v)
{
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

// Reads the header into info / levels (up to max_levels) and, when boxes is
// non-null, the box list of level want_boxes.
int read_header(const char *dir, bx_chk_info *info, bx_chk_level_info *levels,
                int max_levels, int want_boxes, bx_box *boxes, int nbox_cap,
                std::string *err)
{
    std::ifstream H(std::string(dir) + "/Header");
    if (!H)
    {
        *err = std::string("cannot read ") + dir + "/Header";
        return BX_ERR_INVALID;
    }
    std::string tag;
    H >> tag;
    if (tag != CHK_TAG)
    {
        *err = std::string(dir) + "/Header is not a " + CHK_TAG + " header";
        return BX_ERR_INVALID;
    }
    bx_chk_info in{};
    std::string word;
    H >> word >> in.dim >> word >> in.nlev >> word >> in.time >> word >> in.ref_ratio;
    if (!H || in.dim < 1 || in.dim > 3 || in.nlev < 1 || in.nlev > 16)
    {
        *err = "malformed checkpoint header";
        return BX_ERR_INVALID;
    }
    H >> word;
    for (int d = 0; d < 3; ++d)
        H >> in.prob_lo[d];
    H >> word;
    for (int d = 0; d < 3; ++d)
        H >> in.prob_hi[d];
    for (int l = 0; l < in.nlev; ++l)
    {
        bx_chk_level_info li{};
        int               lev = -1;
        H >> word >> lev >> word >> li.level_steps >> word >> li.dt;
        H >> word;
        for (int d = 0; d < 3; ++d)
            H >> li.domain_lo[d];
        for (int d = 0; d < 3; ++d)
            H >> li.domain_hi[d];
        H >> word;
        for (int d = 0; d < 3; ++d)
            H >> li.dx[d];
        H >> word >> li.ng >> word >> li.nbox;
        if (!H || lev != l || li.nbox < 1 || li.ng < 0)
        {
            *err = "malformed level " + std::to_string(l) + " in the checkpoint header";
            return BX_ERR_INVALID;
        }
        for (int b = 0; b < li.nbox; ++b)
        {
            bx_box bx{};
            for (int d = 0; d < 3; ++d)
                H >> bx.lo[d];
            for (int d = 0; d < 3; ++d)
                H >> bx.hi[d];
            if (boxes && l == want_boxes && b < nbox_cap)
                boxes[b] = bx;
        }
        std::string hex;
        H >> word >> li.ndoubles >> word >> hex;
        if (!H)
        {
            *err = "malformed level " + std::to_string(l) + " in the checkpoint header";
            return BX_ERR_INVALID;
        }
        li.checksum = std::strtoull(hex.c_str(), nullptr, 16);
        if (levels && l < max_levels)
            levels[l] = li;
    }
    if (info)
        *info = in;
    return BX_OK;
}

} // namespace

extern "C" int bx_write_checkpoint(const char *dir, int dim, int nlev,
                                   const bx_chk_level *levels, int ref_ratio,
                                   const double prob_lo[3], const double prob_hi[3],
                                   double time, void *stream)
{
    if (!dir || !levels || dim < 1 || dim > 3 || nlev < 1 || nlev > 16 || !prob_lo
        || !prob_hi)
        return set_error(BX_ERR_INVALID, "bx_write_checkpoint: bad argument");
    for (int l = 0; l < nlev; ++l)
        if (!levels[l].boxes || levels[l].nbox < 1 || !levels[l].U || levels[l].ng < 0)
            return set_error(BX_ERR_INVALID, "bx_write_checkpoint: bad level "
                                                 + std::to_string(l));
    const std::string root(dir);
    if (!make_dir(root))
        return set_error(BX_ERR_INVALID, "bx_write_checkpoint: cannot create " + root);
    hipStream_t         s = (hipStream_t)stream;
    std::vector<double> buf;
    std::ostringstream  H;
    H << CHK_TAG << "\n";
    H << "dim " << dim << "\nlevels " << nlev << "\ntime " << fmt(time) << "\nref_ratio "
      << ref_ratio << "\n";
    H << "prob_lo";
    for (int d = 0; d < 3; ++d)
        H << " " << fmt(prob_lo[d]);
    H << "\nprob_hi";
    for (int d = 0; d < 3; ++d)
        H << " " << fmt(prob_hi[d]);
    H << "\n";
    for (int l = 0; l < nlev; ++l)
    {
        const bx_chk_level &L = levels[l];
        int64_t             total = 0;
        if (bx_pack_offsets(dim, L.ng, L.boxes, L.nbox, nullptr, &total) != BX_OK)
            return set_error(BX_ERR_INVALID, "bx_write_checkpoint: bad boxes on level "
                                                 + std::to_string(l));
        const std::string ldir = root + "/Level_" + std::to_string(l);
        if (!make_dir(ldir))
            return set_error(BX_ERR_INVALID, "bx_write_checkpoint: cannot create " + ldir);
        std::ofstream D(ldir + "/State", std::ios::binary | std::ios::trunc);
        if (!D)
            return set_error(BX_ERR_INVALID, "bx_write_checkpoint: cannot write " + ldir);
        uint64_t h = FNV_OFFSET;
        buf.resize((size_t)std::min<int64_t>(total, (int64_t)CHUNK));
        for (int64_t off = 0; off < total; off += (int64_t)CHUNK)
        {
            const size_t n = (size_t)std::min<int64_t>(total - off, (int64_t)CHUNK);
            BX_HIP_CHECK(hipMemcpyAsync(buf.data(), L.U + off, n * sizeof(double),
                                        hipMemcpyDeviceToHost, s));
            BX_HIP_CHECK(hipStreamSynchronize(s));
            h = fnv1a(h, buf.data(), n * sizeof(double));
            D.write((const char *)buf.data(), (std::streamsize)(n * sizeof(double)));
        }
        D.close();
        if (!D)
            return set_error(BX_ERR_INVALID, "bx_write_checkpoint: write failed in " + ldir);
        H << "level " << l << " steps " << L.level_steps << " dt " << fmt(L.dt) << "\n";
        H << "domain";
        for (int d = 0; d < 3; ++d)
            H << " " << L.domain_lo[d];
        for (int d = 0; d < 3; ++d)
            H << " " << L.domain_hi[d];
        H << "\ndx";
        for (int d = 0; d < 3; ++d)
            H << " " << fmt(L.dx[d]);
        H << "\nng " << L.ng << " boxes " << L.nbox << "\n";
        for (int b = 0; b < L.nbox; ++b)
        {
            const bx_box &bx = L.boxes[b];
            H << bx.lo[0] << " " << bx.lo[1] << " " << bx.lo[2] << " " << bx.hi[0] << " "
              << bx.hi[1] << " " << bx.hi[2] << "\n";
        }
        char hex[32];
        std::snprintf(hex, sizeof hex, "%016" PRIx64, h);
        H << "doubles " << total << " fnv1a " << hex << "\n";
    }
    // the header last: a directory with a header holds complete level files
    std::ofstream Hf(root + "/Header", std::ios::trunc);
    Hf << H.str();
    Hf.close();
    if (!Hf)
        return set_error(BX_ERR_INVALID, "bx_write_checkpoint: cannot write " + root + "/Header");
    return BX_OK;
}

extern "C" int bx_read_checkpoint_info(const char *dir, bx_chk_info *info,
                                       bx_chk_level_info *levels, int max_levels)
{
    if (!dir || !info || max_levels < 0 || (max_levels > 0 && !levels))
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_info: bad argument");
    std::string err;
    const int   rc = read_header(dir, info, levels, max_levels, -1, nullptr, 0, &err);
    if (rc != BX_OK)
        return set_error(rc, "bx_read_checkpoint_info: " + err);
    return BX_OK;
}

extern "C" int bx_read_checkpoint_boxes(const char *dir, int level, bx_box *boxes, int nbox)
{
    if (!dir || !boxes || level < 0 || nbox < 1)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_boxes: bad argument");
    bx_chk_info       info;
    bx_chk_level_info li[16];
    std::string       err;
    int rc = read_header(dir, &info, li, 16, level, boxes, nbox, &err);
    if (rc != BX_OK)
        return set_error(rc, "bx_read_checkpoint_boxes: " + err);
    if (level >= info.nlev || li[level].nbox != nbox)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_boxes: level " + std::to_string(level)
                                             + " has " + std::to_string(level < info.nlev ? li[level].nbox : 0)
                                             + " boxes, not " + std::to_string(nbox));
    return BX_OK;
}

extern "C" int bx_read_checkpoint_level(const char *dir, int level, double *U,
                                        int64_t ndoubles, void *stream)
{
    if (!dir || !U || level < 0 || ndoubles < 1)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_level: bad argument");
    bx_chk_info       info;
    bx_chk_level_info li[16];
    std::string       err;
    int               rc = read_header(dir, &info, li, 16, -1, nullptr, 0, &err);
    if (rc != BX_OK)
        return set_error(rc, "bx_read_checkpoint_level: " + err);
    if (level >= info.nlev)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_level: no level " + std::to_string(level));
    if (li[level].ndoubles != ndoubles)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_level: level holds "
                                             + std::to_string(li[level].ndoubles)
                                             + " doubles, buffer " + std::to_string(ndoubles));
    const std::string path = std::string(dir) + "/Level_" + std::to_string(level) + "/State";
    std::ifstream     D(path, std::ios::binary);
    if (!D)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_level: cannot read " + path);
    hipStream_t         s = (hipStream_t)stream;
    std::vector<double> buf((size_t)std::min<int64_t>(ndoubles, (int64_t)CHUNK));
    uint64_t            h = FNV_OFFSET;
    for (int64_t off = 0; off < ndoubles; off += (int64_t)CHUNK)
    {
        const size_t n = (size_t)std::min<int64_t>(ndoubles - off, (int64_t)CHUNK);
        D.read((char *)buf.data(), (std::streamsize)(n * sizeof(double)));
        if (!D)
            return set_error(BX_ERR_INVALID, "bx_read_checkpoint_level: " + path + " is short");
        h = fnv1a(h, buf.data(), n * sizeof(double));
        BX_HIP_CHECK(hipMemcpyAsync(U + off, buf.data(), n * sizeof(double),
                                    hipMemcpyHostToDevice, s));
        BX_HIP_CHECK(hipStreamSynchronize(s));
    }
    if (h != li[level].checksum)
        return set_error(BX_ERR_INVALID, "bx_read_checkpoint_level: checksum mismatch in " + path);
    return BX_OK;
}
