// snapshot.hip — host-side adapter for GOTPM particle files (jread/jwrite format) and the
// 32-byte GOTPM+XYZDBL particle records. No device code: these functions run without a GPU.
//
// File layout (Cosmos/CosmosRW/kjhrw.c:51-220, Params/header.c:28-105, Params/params.h:249,500-518):
//   1. an ASCII header: lines "define <key> = <value> # comment", comment lines starting with '#',
//      ended by the line "#End of Ascii Header\n";
//   2. DM_NP particle records (PMTYPE: dmparticletype, 32 bytes; TREETYPE: treedmparticletype,
//      40 bytes with a leading pointer), then the SPH, star and AGN arrays (empty in GOTPM runs);
//   3. an optional binary dump of the writer's SimParameters struct (kjhrw.c:117,209).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lagpm_internal.h"

namespace {

struct dm32 {  // dmparticletype under GOTPM, eunha.h:389-399
    uint64_t u4if;
    float x, y, z, vx, vy, vz;
};
static_assert(sizeof(dm32) == 32, "GOTPM particle is 32 bytes");
struct dm40 {  // treedmparticletype under GOTPM, eunha.h:400-411
    uint64_t next;
    uint64_t u4if;
    float x, y, z, vx, vy, vz;
};
static_assert(sizeof(dm40) == 40, "GOTPM tree particle is 40 bytes");

constexpr const char* kClosing = "#End of Ascii Header\n";  // P_Closing, params.h:249

struct file_closer {
    FILE* f;
    ~file_closer() {
        if (f) fclose(f);
    }
};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\n' || s[b - 1] == '\r')) b--;
    return s.substr(a, b - a);
}

// One header line "define <key> = <value> [# comment]": the key is the text before the first '='
// (it may hold spaces, "Particle Type"); the value is the first whitespace-delimited token after
// it, as the reference's sscanf formats read it (S_INT/S_LONG/S_FLOAT/S_DOUBLE/S_STRING).
bool parse_define(const std::string& line, std::string& key, std::string& val) {
    if (line.compare(0, 7, "define ") != 0) return false;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) return false;
    key = trim(line.substr(7, eq - 7));
    size_t a = eq + 1;
    while (a < line.size() && (line[a] == ' ' || line[a] == '\t')) a++;
    size_t b = a;
    while (b < line.size() && line[b] != ' ' && line[b] != '\t' && line[b] != '\n' && line[b] != '\r') b++;
    val = line.substr(a, b - a);
    return true;
}

// Reads the header lines up to and including the closing line. Returns false when the file ends
// first. `lines` receives every line as read (byte for byte), for a writer that keeps them.
bool read_header_lines(FILE* f, std::vector<std::string>& lines) {
    std::string cur;
    int c;
    while ((c = fgetc(f)) != EOF) {
        cur.push_back((char)c);
        if (c == '\n') {
            lines.push_back(cur);
            if (cur == kClosing) return true;
            cur.clear();
        }
    }
    return false;
}

int fail(lagpm_ctx* ctx, int code, const std::string& msg) { return lagpm_fail(ctx, code, msg); }

bool to_long(const std::string& v, long& out) {
    if (v.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const long r = strtol(v.c_str(), &end, 10);
    if (errno || !end || *end) return false;
    out = r;
    return true;
}
bool to_double(const std::string& v, double& out) {
    if (v.empty()) return false;
    char* end = nullptr;
    errno = 0;
    const double r = strtod(v.c_str(), &end);
    if (errno || !end || *end) return false;
    out = r;
    return true;
}

// the header of `path` into info; the particle arrays start at info->header_bytes
int read_info(lagpm_ctx* ctx, const char* path, lagpm_snapshot_info* info, std::vector<std::string>* keep) {
    if (!path || !info) return fail(ctx, LAGPM_E_INVAL, "NULL argument");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(ctx, LAGPM_E_IO, std::string("cannot open ") + path + ": " + strerror(errno));
    file_closer fc{f};
    std::vector<std::string> lines;
    if (!read_header_lines(f, lines)) return fail(ctx, LAGPM_E_IO, std::string(path) + ": no \"#End of Ascii Header\" line");
    lagpm_snapshot_info s;
    memset(&s, 0, sizeof(s));
    size_t bytes = 0;
    int bad = 0;
    std::string badkey;
    struct lfield { const char* key; long* dst; };
    struct dfield { const char* key; double* dst; };
    long ptype = 0, nspace = 1, stepcount = 0, xyzshift = 0;
    const lfield lf[] = {{"Nx", &s.nx},           {"Ny", &s.ny},           {"Nz", &s.nz},
                         {"Np", &s.np_sum},       {"DNp", &s.np_dm},       {"BNp", &s.np_sph},
                         {"SNp", &s.np_star},     {"ANp", &s.np_agn},      {"Nspace", &nspace},
                         {"Particle Type", &ptype}, {"Stepcount", &stepcount}, {"XYZ_Shift_Flag", &xyzshift}};
    const dfield df[] = {{"Amax", &s.amax},       {"Anow", &s.anow},         {"Astep", &s.astep},
                         {"OmegaMatter0", &s.omep}, {"OmegaLambda0", &s.omeplam}, {"Hubble", &s.hubble},
                         {"Boxsize(Mpc/h)", &s.boxsize}, {"Theta", &s.theta}, {"EPSILON", &s.epsilon},
                         {"Rsphere", &s.rsphere}};
    for (const std::string& ln : lines) {
        bytes += ln.size();
        std::string k, v;
        if (ln.empty() || ln[0] == '#' || !parse_define(ln, k, v)) continue;
        bool ok = true;
        for (const lfield& q : lf)
            if (k == q.key) ok = to_long(v, *q.dst);
        for (const dfield& q : df)
            if (k == q.key) ok = to_double(v, *q.dst);
        if (!ok && !bad) {
            bad = 1;
            badkey = k + " = '" + v + "'";
        }
    }
    s.ptype = (int)ptype;
    s.nspace = (int)nspace;
    s.stepcount = (int)stepcount;
    s.xyzshift = (int)xyzshift;
    if (bad) return fail(ctx, LAGPM_E_IO, std::string(path) + ": unreadable header value " + badkey);
    if (s.ptype != 0 && s.ptype != 1)
        return fail(ctx, LAGPM_E_IO, std::string(path) + ": unknown particle type " + std::to_string(s.ptype));
    if (s.np_dm < 0 || s.np_sph < 0 || s.np_star < 0 || s.np_agn < 0)
        return fail(ctx, LAGPM_E_IO, std::string(path) + ": negative particle count");
    s.header_bytes = bytes;
    s.record_bytes = s.ptype == 1 ? sizeof(dm40) : sizeof(dm32);
    if (fseek(f, 0, SEEK_END) != 0) return fail(ctx, LAGPM_E_IO, std::string(path) + ": cannot seek");
    const long end = ftell(f);
    if (end < 0) return fail(ctx, LAGPM_E_IO, std::string(path) + ": cannot tell size");
    s.file_bytes = (size_t)end;
    const size_t need = bytes + (size_t)s.np_dm * s.record_bytes;
    if (s.file_bytes < need)
        return fail(ctx, LAGPM_E_IO, std::string(path) + ": file ends inside the DM particle array (" +
                                         std::to_string(s.file_bytes) + " < " + std::to_string(need) + " bytes)");
    *info = s;
    if (keep) *keep = std::move(lines);
    return LAGPM_OK;
}

}  // namespace

extern "C" {

int lagpm_snapshot_read_header(lagpm_ctx* ctx, const char* path, lagpm_snapshot_info* info) {
    return read_info(ctx, path, info, nullptr);
}

int lagpm_snapshot_read_dm(lagpm_ctx* ctx, const char* path, void* dm_bp, size_t np) {
    lagpm_snapshot_info s;
    RCHK(read_info(ctx, path, &s, nullptr));
    if (np != (size_t)s.np_dm)
        return fail(ctx, LAGPM_E_INVAL, "np " + std::to_string(np) + " differs from the file's DNp " + std::to_string(s.np_dm));
    if (np && !dm_bp) return fail(ctx, LAGPM_E_INVAL, "NULL record buffer");
    FILE* f = fopen(path, "rb");
    if (!f) return fail(ctx, LAGPM_E_IO, std::string("cannot open ") + path);
    file_closer fc{f};
    if (fseek(f, (long)s.header_bytes, SEEK_SET) != 0) return fail(ctx, LAGPM_E_IO, "cannot seek to the particles");
    dm32* out = (dm32*)dm_bp;
    if (s.ptype == 0) {
        if (np && fread(out, sizeof(dm32), np, f) != np) return fail(ctx, LAGPM_E_IO, "short read of DM particles");
    } else {
        // tree records: the leading linked-list pointer is dropped (kjhrw.c:103-107)
        const size_t chunk = 1 << 16;
        std::vector<dm40> buf(std::min(np, chunk));
        for (size_t done = 0; done < np;) {
            const size_t m = std::min(chunk, np - done);
            if (fread(buf.data(), sizeof(dm40), m, f) != m) return fail(ctx, LAGPM_E_IO, "short read of DM tree particles");
            for (size_t i = 0; i < m; i++) {
                const dm40& q = buf[i];
                out[done + i] = dm32{q.u4if, q.x, q.y, q.z, q.vx, q.vy, q.vz};
            }
            done += m;
        }
    }
    return LAGPM_OK;
}

int lagpm_snapshot_write(lagpm_ctx* ctx, const char* src_path, const char* dst_path, const void* dm_bp, size_t np,
                         double anow, int stepcount) {
    if (!src_path || !dst_path || (np && !dm_bp)) return fail(ctx, LAGPM_E_INVAL, "NULL argument");
    if (!strcmp(src_path, dst_path)) return fail(ctx, LAGPM_E_INVAL, "the output would overwrite its template");
    lagpm_snapshot_info s;
    std::vector<std::string> lines;
    RCHK(read_info(ctx, src_path, &s, &lines));
    if (s.np_sph || s.np_star || s.np_agn)
        return fail(ctx, LAGPM_E_INVAL, "template holds SPH/star/AGN particles: only DM-only (GOTPM) files are written");
    if (np != (size_t)s.np_dm)
        return fail(ctx, LAGPM_E_INVAL, "np " + std::to_string(np) + " differs from the template's DNp " + std::to_string(s.np_dm));
    // No trailing SimParameters block is written: jread freads one over the parsed header after
    // the particles (Cosmos/CosmosRW/kjhrw.c:118), so a block carried over from the template would
    // put the template's Anow, Stepcount and particle type back; without it that fread reads
    // nothing and the header written here stands.
    FILE* out = fopen(dst_path, "wb");
    if (!out) return fail(ctx, LAGPM_E_IO, std::string("cannot create ") + dst_path + ": " + strerror(errno));
    file_closer fo{out};
    // header: the template's lines with the time lines rewritten in the reference's formats
    // (P_Anow "%lg", P_Stepcount "%d", params.h:64,80) and the particle type set to PMTYPE
    bool ok = true;
    for (const std::string& ln : lines) {
        std::string k, v;
        if (ln[0] != '#' && parse_define(ln, k, v)) {
            char buf[256];
            const size_t eq = ln.find('=');
            const std::string head = ln.substr(0, eq + 1);
            if (k == "Anow") {
                snprintf(buf, sizeof(buf), " %.17g # Current expansion factor/ current time\n", anow);
                ok = ok && fputs((head + buf).c_str(), out) >= 0;
                continue;
            }
            if (k == "Stepcount" || k == "Stepnum") {
                snprintf(buf, sizeof(buf), " %d\n", stepcount);
                ok = ok && fputs((head + buf).c_str(), out) >= 0;
                continue;
            }
            if (k == "Particle Type") {
                ok = ok && fputs((head + " 0\n").c_str(), out) >= 0;
                continue;
            }
        }
        ok = ok && fwrite(ln.data(), 1, ln.size(), out) == ln.size();
    }
    if (np) ok = ok && fwrite(dm_bp, sizeof(dm32), np, out) == np;
    if (!ok) return fail(ctx, LAGPM_E_IO, std::string("write to ") + dst_path + " failed");
    return LAGPM_OK;
}

int lagpm_eunha_decode(const void* dm_bp, size_t np, long nx, double* x, double* y, double* z, double* vx, double* vy,
                       double* vz) {
    if (nx <= 0 || (np && (!dm_bp || !x || !y || !z))) return LAGPM_E_INVAL;
    const dm32* p = (const dm32*)dm_bp;
    const double L = (double)nx;
    const uint64_t n1 = (uint64_t)nx, n2 = n1 * n1;
    for (size_t i = 0; i < np; i++) {
        const uint64_t idx = (p[i].u4if << 8) >> 8;  // PINDX: the low 56 bits, eunha.h:1212
        const double ii = (double)(idx % n1), jj = (double)((idx % n2) / n1), kk = (double)(idx / n2);
        // XofP/YofP/ZofP, eunha.h:1228-1232
        x[i] = std::fmod((double)p[i].x + ii + L, L);
        y[i] = std::fmod((double)p[i].y + jj + L, L);
        z[i] = std::fmod((double)p[i].z + kk + L, L);
        if (vx) vx[i] = p[i].vx;
        if (vy) vy[i] = p[i].vy;
        if (vz) vz[i] = p[i].vz;
    }
    return LAGPM_OK;
}

int lagpm_eunha_encode(void* dm_bp, size_t np, long nx, const double* x, const double* y, const double* z,
                       const double* vx, const double* vy, const double* vz) {
    if (nx <= 0 || (np && !dm_bp)) return LAGPM_E_INVAL;
    dm32* p = (dm32*)dm_bp;
    const double L = (double)nx, h = 0.5 * L;
    const uint64_t n1 = (uint64_t)nx, n2 = n1 * n1;
    // the displacement from the record's own lattice site, taken as the periodic image nearest to
    // it ([-n/2, n/2)), so XofP of the record gives the position back
    auto disp = [&](double pos, double site) {
        double d = std::fmod(pos - site, L);
        if (d >= h) d -= L;
        if (d < -h) d += L;
        return (float)d;
    };
    for (size_t i = 0; i < np; i++) {
        const uint64_t idx = (p[i].u4if << 8) >> 8;
        const double ii = (double)(idx % n1), jj = (double)((idx % n2) / n1), kk = (double)(idx / n2);
        if (x) p[i].x = disp(x[i], ii);
        if (y) p[i].y = disp(y[i], jj);
        if (z) p[i].z = disp(z[i], kk);
        if (vx) p[i].vx = (float)vx[i];
        if (vy) p[i].vy = (float)vy[i];
        if (vz) p[i].vz = (float)vz[i];
    }
    return LAGPM_OK;
}

}  // extern "C"
