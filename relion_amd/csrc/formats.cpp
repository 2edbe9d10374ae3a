// STAR metadata tables and MRC images (host code): the file formats the E-step's inputs come from and its maps go
// to. Follows the reference's readers and writers (src/metadata_table.cpp, src/strings.cpp, src/rwMRC.h,
// src/image.h) in what they accept, refuse and produce; the data model is our own (values kept as text per cell).
#include <fcntl.h>
#include <omp.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <string>
#include <vector>

#include "../../include/rln_formats.h"
#include "rln_common.h"

#define RLN_TRY(expr)                  \
    do {                               \
        int _rc = (expr);              \
        if (_rc != RLN_OK) return _rc; \
    } while (0)

struct rln_star {
    struct Block {
        std::string name;
        bool is_list = false;
        std::vector<std::string> labels;
        std::vector<std::string> cells;  // row-major [rows][labels]
        long rows = 0;
    };
    int version = 30000;
    std::vector<Block> blocks;
};

namespace {

constexpr int kCurrentStarVersion = 50001;  // CURRENT_MDT_VERSION, src/metadata_table.h:49

// unescape + simplify (src/strings.cpp:64-86, 122-160): control characters dropped, tabs to spaces, leading spaces
// removed, runs of spaces collapsed to one.
std::string simplify(const std::string &s) {
    std::string u;
    u.reserve(s.size());
    for (char c : s) {
        if (c == '\t')
            u += ' ';
        else if (c != '\n' && c != '\v' && c != '\b' && c != '\r' && c != '\f' && c != '\a')
            u += c;
    }
    std::string out;
    out.reserve(u.size());
    size_t i = u.find_first_not_of(' ');
    if (i == std::string::npos) return out;
    bool prev_space = false;
    for (; i < u.size(); ++i) {
        const bool sp = u[i] == ' ';
        if (sp && prev_space) continue;
        out += u[i];
        prev_space = sp;
    }
    return out;
}

std::string trim(const std::string &s) {
    const size_t a = s.find_first_not_of(" \t\n\r\f\v");
    if (a == std::string::npos) return std::string();
    const size_t b = s.find_last_not_of(" \t\n\r\f\v");
    return s.substr(a, b - a + 1);
}

// nextTokenInSTAR (src/strings.cpp:595-662): whitespace-separated words, '#' starts a comment, single or double
// quotes enclose a word that ends at a closing quote followed by whitespace or the end; '\a' escapes a quote.
// Returns 1 with a token, 0 at the end, -1 for an unclosed quote.
int next_token(const std::string &str, size_t &i, std::string &tok) {
    const size_t len = str.size();
    if (i >= len) return 0;
    size_t start = str.find_first_not_of(" \t\n", i);
    if (start == std::string::npos) return 0;
    if (str[start] == '#') return 0;
    tok.clear();
    if (str[start] == '\'' || str[start] == '"') {
        const char quote = str[start];
        ++start;
        size_t pos = start;
        bool closed = false;
        while (pos < len) {
            if (str[pos] == quote && str[pos - 1] != '\a') {
                const size_t nx = pos + 1;
                if (nx == len || str[nx] == ' ' || str[nx] == '\t' || str[nx] == '\n') {
                    closed = true;
                    break;
                }
            }
            if (str[pos] != '\a') tok += str[pos];
            ++pos;
        }
        if (!closed) return -1;
        i = pos + 1;
    } else {
        size_t end = str.find_first_of(" \t\n", start + 1);
        if (end == std::string::npos) end = len;
        tok = str.substr(start, end - start);
        i = end + 1;
    }
    return 1;
}

// escapeStringForSTAR (src/strings.cpp:88-118).
std::string escape(const std::string &v) {
    if (v.empty()) return "\"\"";
    if (v[0] != '"' && v[0] != '\'' && v.find_first_of(" \t") == std::string::npos) return v;
    std::string e = "\"";
    for (size_t p = 0; p < v.size(); ++p) {
        if (v[p] == '"') {
            const size_t nx = p + 1;
            if (nx == v.size() - 1 || (nx < v.size() && (v[nx] == ' ' || v[nx] == '\t'))) e += '\a';
        }
        e += v[p];
    }
    return e + "\"";
}

// Label of a "_label #n" line: the first word after the underscore (readStarLoop, src/metadata_table.cpp:1052-1061).
std::string loop_label(const std::string &line) {
    size_t a = line.find('_') + 1;
    size_t b = line.find_first_of(" #", a);
    return line.substr(a, b == std::string::npos ? std::string::npos : b - a);
}

// readStarLoop (src/metadata_table.cpp:1036-1130). lines[i] is the line after "loop_"; returns the next line.
int parse_loop(const std::vector<std::string> &lines, size_t &i, rln_star::Block &b) {
    b.is_list = false;
    std::string line;
    bool have_row = false;
    while (i < lines.size()) {
        line = simplify(lines[i++]);
        if (line.empty() || line[0] == '#' || line[0] == ';') continue;
        if (line[0] == '_') {
            b.labels.push_back(loop_label(line));
            continue;
        }
        have_row = true;
        break;
    }
    const size_t nl = b.labels.size();
    while (have_row) {
        if (line.empty()) break;  // an empty line ends the table
        std::vector<std::string> vals;
        size_t pos = 0;
        std::string tok;
        int r;
        while ((r = next_token(line, pos, tok)) == 1) {
            if (vals.size() >= nl)
                RLN_FAIL(RLN_ERR_ARG, "A line in the STAR file contains more columns than the number of labels: %s",
                         line.c_str());
            vals.push_back(tok);
        }
        if (r < 0) RLN_FAIL(RLN_ERR_ARG, "Could not find closing quote in a STAR file line: %s", line.c_str());
        if (vals.size() < nl && nl > 2)
            RLN_FAIL(RLN_ERR_ARG,
                     "A line in the STAR file contains fewer columns than the number of labels. Expected = %zu "
                     "Found = %zu",
                     nl, vals.size());
        vals.resize(nl);
        b.cells.insert(b.cells.end(), vals.begin(), vals.end());
        ++b.rows;
        if (i >= lines.size()) break;
        line = simplify(lines[i++]);
    }
    return RLN_OK;
}

// readStarList (src/metadata_table.cpp:1132-1201). Sets *loop_follows when a loop_ line ends the list.
int parse_list(const std::vector<std::string> &lines, size_t &i, rln_star::Block &b, bool *loop_follows) {
    b.is_list = true;
    b.rows = 1;
    *loop_follows = false;
    std::vector<std::string> vals;
    while (i < lines.size()) {
        const std::string &line = lines[i++];
        size_t pos = 0;
        std::string first, value;
        const int r = next_token(line, pos, first);
        if (r < 0) RLN_FAIL(RLN_ERR_ARG, "Could not find closing quote in a STAR file line: %s", line.c_str());
        if (r == 0) continue;
        if (first[0] == '_') {
            const int rv = next_token(line, pos, value);
            if (rv <= 0)
                RLN_FAIL(RLN_ERR_ARG, "MetaDataTable::readStarList: did not encounter a single word after %s",
                         first.c_str());
            b.labels.push_back(first.substr(1));
            vals.push_back(value);
        } else if (first[0] == '#' || first[0] == ';') {
            continue;
        } else if (first.compare(0, 5, "loop_") == 0) {
            *loop_follows = true;
            break;
        } else if (first.compare(0, 5, "data_") == 0) {
            --i;  // the next block starts here
            break;
        }
    }
    b.cells = vals;
    return RLN_OK;
}

// MetaDataTable::readAll (src/metadata_table.cpp:1275-1345).
int parse_star(const char *text, size_t len, rln_star &s) {
    std::vector<std::string> lines;
    size_t a = 0;
    while (a < len) {
        const char *nl = (const char *)memchr(text + a, '\n', len - a);
        const size_t b = nl ? (size_t)(nl - text) : len;
        lines.emplace_back(text + a, b - a);
        a = b + 1;
    }
    for (const std::string &l : lines)
        if (l.size() >= 2 && l.back() == '\r')
            RLN_FAIL(RLN_ERR_ARG,
                     "RELION does not support CR+LF as a new line code (convert the STAR file to LF line ends)");
    s.version = 30000;
    size_t i = 0;
    while (i < lines.size()) {
        const std::string line = trim(lines[i++]);
        const size_t vp = line.find("# version ");
        if (vp != std::string::npos) s.version = std::atoi(line.c_str() + vp + 10);
        const size_t dp = line.find("data_");
        if (dp == std::string::npos) continue;
        rln_star::Block b;
        b.name = line.substr(dp + 5);
        while (i < lines.size()) {
            const std::string &l = lines[i];
            if (l.find("loop_") != std::string::npos) {
                ++i;
                RLN_TRY(parse_loop(lines, i, b));
                break;
            }
            if (!l.empty() && l[0] == '_') {
                bool loop_follows = false;
                RLN_TRY(parse_list(lines, i, b, &loop_follows));
                if (loop_follows) {  // the reference keeps the list only: skip the loop's labels and rows
                    while (i < lines.size() && !simplify(lines[i]).empty() && simplify(lines[i])[0] == '_') ++i;
                    while (i < lines.size() && !simplify(lines[i]).empty()) ++i;
                }
                break;
            }
            if (trim(l).find("data_") != std::string::npos) break;  // an empty block
            ++i;
        }
        s.blocks.push_back(std::move(b));
    }
    return RLN_OK;
}

bool valid_block(const rln_star *s, int b) { return s && b >= 0 && b < (int)s->blocks.size(); }

// ------------------------------------------------------------------------------------------------ MRC
constexpr int kMrcHeader = 1024;
constexpr int kSwapTrig = 65535;  // SWAPTRIG, src/image.h:386

struct MrcHeader {  // the 1024-byte header (MRC2014), fields used here
    int32_t nx, ny, nz, mode, nxstart, nystart, nzstart, mx, my, mz;
    float a, b, c, alpha, beta, gamma;
    int32_t mapc, mapr, maps;
    float amin, amax, amean;
    int32_t ispg, nsymbt;
    float extra[25];
    float xorigin, yorigin, zorigin;
    char map[4];
    unsigned char machst[4];
    float arms;
    int32_t nlabl;
    char labels[800];
};
static_assert(sizeof(MrcHeader) == kMrcHeader, "MRC header layout");

void swap4(void *p) {
    unsigned char *c = (unsigned char *)p;
    std::swap(c[0], c[3]);
    std::swap(c[1], c[2]);
}

int mrc_bits(int mode) {
    switch (mode) {
        case 0: return 8;
        case 1: return 16;
        case 2: return 32;
        case 6: return 16;
        case 12: return 16;
        case 101: return 4;
        default: return 0;
    }
}

// IEEE 754 binary16 to binary32 from the format's definition (sign, 5-bit exponent with bias 15, 10-bit fraction;
// subnormals, infinities and NaNs).
float half_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    const int e = (h >> 10) & 31;
    const uint32_t f = h & 1023;
    uint32_t bits;
    if (e == 0) {
        if (f == 0) {
            bits = sign;
        } else {  // subnormal: f * 2^-24
            float v = std::ldexp((float)f, -24);
            std::memcpy(&bits, &v, 4);
            bits |= sign;
        }
    } else if (e == 31) {
        bits = sign | 0x7f800000u | (f << 13);
    } else {
        bits = sign | ((uint32_t)(e - 15 + 127) << 23) | (f << 13);
    }
    float out;
    std::memcpy(&out, &bits, 4);
    return out;
}

int read_header(int fd, const char *path, MrcHeader &h, rln_mrc_info &info) {
    if (pread(fd, &h, kMrcHeader, 0) != kMrcHeader) RLN_FAIL(RLN_ERR_ARG, "rwMRC: error in reading header of image %s", path);
    info = rln_mrc_info{};
    if (std::abs(h.mode) > kSwapTrig || std::abs(h.nx) > kSwapTrig) {  // parseMRCHeader, src/rwMRC.h:149-157
        info.swapped = 1;
        unsigned char *b = (unsigned char *)&h;
        for (int k = 0; k < kMrcHeader - 800; k += 4) swap4(b + k);
    }
    info.nx = h.nx;
    info.ny = h.ny;
    info.nz = h.nz;
    info.mode = h.mode;
    info.nsymbt = h.nsymbt;
    if (h.mode == 3 || h.mode == 4) RLN_FAIL(RLN_ERR_ARG, "readMRC: only real-space images may be read into RELION (%s)", path);
    if (!mrc_bits(h.mode)) RLN_FAIL(RLN_ERR_ARG, "readMRC: unsupported MRC mode %d (%s)", h.mode, path);
    if (h.mode == 101 && ((long long)h.nx * h.ny) % 2)
        RLN_FAIL(RLN_ERR_ARG, "Currently we support 4-bit MRC (mode 101) only when nx * ny is an even number.");
    if (h.nx <= 0 || h.ny <= 0 || h.nz <= 0 || h.nsymbt < 0)
        RLN_FAIL(RLN_ERR_ARG, "MRC header of %s: invalid sizes %d x %d x %d (nsymbt %d)", path, h.nx, h.ny, h.nz, h.nsymbt);
    info.pixel_size = (h.mx && h.a != 0.f) ? (double)h.a / h.mx : 0.;
    info.pixel_size_z = (h.mz && h.c != 0.f) ? (double)h.c / h.mz : 0.;
    info.amin = h.amin;
    info.amax = h.amax;
    info.amean = h.amean;
    info.arms = h.arms;
    return RLN_OK;
}

// Reads `count` sections from `first` of an open file into out, converting to float32.
int read_sections(int fd, const char *path, const rln_mrc_info &info, long first, long count, float *out) {
    if (first < 0 || count < 0 || first + count > info.nz)
        RLN_FAIL(RLN_ERR_ARG, "readMRC: images %ld..%ld outside the stack of %d in %s", first + 1, first + count, info.nz,
                 path);
    const size_t px = (size_t)info.nx * info.ny;
    const int bits = mrc_bits(info.mode);
    const size_t sec_bytes = px * bits / 8;
    const off_t off = (off_t)kMrcHeader + info.nsymbt + (off_t)first * (off_t)sec_bytes;
    const size_t n = px * (size_t)count, nbytes = sec_bytes * (size_t)count;
    if (info.mode == 2) {
        char *dst = (char *)out;
        size_t done = 0;
        while (done < nbytes) {
            const ssize_t r = pread(fd, dst + done, nbytes - done, off + (off_t)done);
            if (r <= 0) RLN_FAIL(RLN_ERR_ARG, "readMRC: short read of %s (%s)", path, r < 0 ? strerror(errno) : "end of file");
            done += (size_t)r;
        }
        if (info.swapped)
            for (size_t k = 0; k < n; ++k) swap4(out + k);
        return RLN_OK;
    }
    std::vector<unsigned char> raw(nbytes);
    size_t done = 0;
    while (done < nbytes) {
        const ssize_t r = pread(fd, raw.data() + done, nbytes - done, off + (off_t)done);
        if (r <= 0) RLN_FAIL(RLN_ERR_ARG, "readMRC: short read of %s (%s)", path, r < 0 ? strerror(errno) : "end of file");
        done += (size_t)r;
    }
    const unsigned char *p = raw.data();
    const bool sw = info.swapped != 0;
    auto u16 = [&](size_t k) -> uint16_t {
        const uint16_t v = (uint16_t)(p[2 * k] | (p[2 * k + 1] << 8));
        return sw ? (uint16_t)((v >> 8) | (v << 8)) : v;
    };
    switch (info.mode) {
        case 0:
            for (size_t k = 0; k < n; ++k) out[k] = (float)(signed char)p[k];
            break;
        case 1:
            for (size_t k = 0; k < n; ++k) out[k] = (float)(int16_t)u16(k);
            break;
        case 6:
            for (size_t k = 0; k < n; ++k) out[k] = (float)u16(k);
            break;
        case 12:
            for (size_t k = 0; k < n; ++k) out[k] = half_to_float(u16(k));
            break;
        case 101:
            for (size_t k = 0; k < n / 2; ++k) {
                out[2 * k] = (float)(p[k] & 15);
                out[2 * k + 1] = (float)((p[k] >> 4) & 15);
            }
            break;
        default:
            RLN_FAIL(RLN_ERR_ARG, "readMRC: unsupported MRC mode %d", info.mode);
    }
    return RLN_OK;
}

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) close(fd);
    }
};

}  // namespace

// ------------------------------------------------------------------------------------------------ STAR C ABI
extern "C" int rln_star_parse(const char *text, size_t len, rln_star **out) {
    if (!out || (!text && len)) RLN_FAIL(RLN_ERR_ARG, "rln_star_parse: null argument");
    rln_star *s = new rln_star();
    const int rc = parse_star(text, len, *s);
    if (rc != RLN_OK) {
        delete s;
        return rc;
    }
    *out = s;
    return RLN_OK;
}

extern "C" int rln_star_read(const char *path, rln_star **out) {
    if (!path || !out) RLN_FAIL(RLN_ERR_ARG, "rln_star_read: null argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) RLN_FAIL(RLN_ERR_ARG, "cannot open STAR file %s: %s", path, strerror(errno));
    std::string text;
    char buf[1 << 16];
    size_t r;
    while ((r = std::fread(buf, 1, sizeof(buf), f)) > 0) text.append(buf, r);
    std::fclose(f);
    return rln_star_parse(text.data(), text.size(), out);
}

extern "C" int rln_star_create(rln_star **out) {
    if (!out) RLN_FAIL(RLN_ERR_ARG, "rln_star_create: null output");
    *out = new rln_star();
    (*out)->version = kCurrentStarVersion;
    return RLN_OK;
}

extern "C" int rln_star_destroy(rln_star *s) {
    delete s;
    return RLN_OK;
}

extern "C" int rln_star_version(const rln_star *s) { return s ? s->version : RLN_ERR_ARG; }
extern "C" int rln_star_block_count(const rln_star *s) { return s ? (int)s->blocks.size() : RLN_ERR_ARG; }
extern "C" const char *rln_star_block_name(const rln_star *s, int b) {
    return valid_block(s, b) ? s->blocks[b].name.c_str() : nullptr;
}
extern "C" int rln_star_find_block(const rln_star *s, const char *name) {
    if (!s || !name) return -1;
    for (size_t b = 0; b < s->blocks.size(); ++b)
        if (s->blocks[b].name == name) return (int)b;
    return -1;
}
extern "C" int rln_star_block_is_list(const rln_star *s, int b) {
    if (!valid_block(s, b)) RLN_FAIL(RLN_ERR_ARG, "no STAR block %d", b);
    return s->blocks[b].is_list ? 1 : 0;
}
extern "C" long rln_star_rows(const rln_star *s, int b) {
    if (!valid_block(s, b)) RLN_FAIL(RLN_ERR_ARG, "no STAR block %d", b);
    return s->blocks[b].rows;
}
extern "C" int rln_star_columns(const rln_star *s, int b) {
    if (!valid_block(s, b)) RLN_FAIL(RLN_ERR_ARG, "no STAR block %d", b);
    return (int)s->blocks[b].labels.size();
}
extern "C" const char *rln_star_label(const rln_star *s, int b, int c) {
    if (!valid_block(s, b) || c < 0 || c >= (int)s->blocks[b].labels.size()) return nullptr;
    return s->blocks[b].labels[c].c_str();
}
extern "C" int rln_star_find_label(const rln_star *s, int b, const char *label) {
    if (!valid_block(s, b) || !label) return -1;
    const auto &L = s->blocks[b].labels;
    for (size_t c = 0; c < L.size(); ++c)
        if (L[c] == label) return (int)c;
    return -1;
}
extern "C" const char *rln_star_value(const rln_star *s, int b, long row, int c) {
    if (!valid_block(s, b)) return nullptr;
    const auto &B = s->blocks[b];
    if (row < 0 || row >= B.rows || c < 0 || c >= (int)B.labels.size()) return nullptr;
    return B.cells[(size_t)row * B.labels.size() + c].c_str();
}

extern "C" int rln_star_column_double(const rln_star *s, int b, int c, double *out) {
    if (!valid_block(s, b) || !out || c < 0 || c >= (int)s->blocks[b].labels.size())
        RLN_FAIL(RLN_ERR_ARG, "rln_star_column_double: no column %d in block %d", c, b);
    const auto &B = s->blocks[b];
    const size_t nl = B.labels.size();
    for (long r = 0; r < B.rows; ++r) {
        const std::string &v = B.cells[(size_t)r * nl + c];
        char *end = nullptr;
        errno = 0;
        const double d = std::strtod(v.c_str(), &end);
        if (v.empty() || *end != '\0')
            RLN_FAIL(RLN_ERR_ARG, "STAR value '%s' of %s (row %ld) is not a number", v.c_str(), B.labels[c].c_str(), r + 1);
        out[r] = d;
    }
    return RLN_OK;
}

extern "C" int rln_star_column_long(const rln_star *s, int b, int c, long *out) {
    if (!valid_block(s, b) || !out || c < 0 || c >= (int)s->blocks[b].labels.size())
        RLN_FAIL(RLN_ERR_ARG, "rln_star_column_long: no column %d in block %d", c, b);
    const auto &B = s->blocks[b];
    const size_t nl = B.labels.size();
    for (long r = 0; r < B.rows; ++r) {
        const std::string &v = B.cells[(size_t)r * nl + c];
        char *end = nullptr;
        const long d = std::strtol(v.c_str(), &end, 10);
        if (v.empty() || *end != '\0')
            RLN_FAIL(RLN_ERR_ARG, "STAR value '%s' of %s (row %ld) is not an integer", v.c_str(), B.labels[c].c_str(), r + 1);
        out[r] = d;
    }
    return RLN_OK;
}

extern "C" int rln_star_add_block(rln_star *s, const char *name, int is_list, int *block) {
    if (!s || !name) RLN_FAIL(RLN_ERR_ARG, "rln_star_add_block: null argument");
    rln_star::Block b;
    b.name = name;
    b.is_list = is_list != 0;
    b.rows = b.is_list ? 1 : 0;
    s->blocks.push_back(std::move(b));
    if (block) *block = (int)s->blocks.size() - 1;
    return RLN_OK;
}

extern "C" int rln_star_add_column(rln_star *s, int b, const char *label, const char *def, int *column) {
    if (!valid_block(s, b) || !label || !*label) RLN_FAIL(RLN_ERR_ARG, "rln_star_add_column: bad block or label");
    auto &B = s->blocks[b];
    if (rln_star_find_label(s, b, label) >= 0) RLN_FAIL(RLN_ERR_ARG, "label %s already in block %s", label, B.name.c_str());
    const size_t nl = B.labels.size();
    std::vector<std::string> cells;
    cells.reserve((size_t)B.rows * (nl + 1));
    for (long r = 0; r < B.rows; ++r) {
        for (size_t c = 0; c < nl; ++c) cells.push_back(std::move(B.cells[(size_t)r * nl + c]));
        cells.push_back(def ? def : "");
    }
    B.cells.swap(cells);
    B.labels.push_back(label);
    if (column) *column = (int)nl;
    return RLN_OK;
}

extern "C" int rln_star_add_rows(rln_star *s, int b, long n) {
    if (!valid_block(s, b) || n < 0) RLN_FAIL(RLN_ERR_ARG, "rln_star_add_rows: bad block or count");
    auto &B = s->blocks[b];
    if (B.is_list && n > 0) RLN_FAIL(RLN_ERR_ARG, "a list block holds one row");
    B.cells.resize(B.cells.size() + (size_t)n * B.labels.size());
    B.rows += n;
    return RLN_OK;
}

extern "C" int rln_star_set_value(rln_star *s, int b, long row, int c, const char *value) {
    if (!valid_block(s, b) || !value) RLN_FAIL(RLN_ERR_ARG, "rln_star_set_value: bad block or value");
    auto &B = s->blocks[b];
    if (row < 0 || row >= B.rows || c < 0 || c >= (int)B.labels.size())
        RLN_FAIL(RLN_ERR_ARG, "rln_star_set_value: row %ld column %d outside block %s", row, c, B.name.c_str());
    B.cells[(size_t)row * B.labels.size() + c] = value;
    return RLN_OK;
}

extern "C" int rln_star_write(const rln_star *s, const char *path) {
    if (!s || !path) RLN_FAIL(RLN_ERR_ARG, "rln_star_write: null argument");
    std::string o;
    for (const auto &B : s->blocks) {
        o += "\n# version " + std::to_string(kCurrentStarVersion) + "\n\ndata_" + B.name + "\n\n";
        const size_t nl = B.labels.size();
        if (!B.is_list) {
            o += "loop_ \n";
            for (size_t c = 0; c < nl; ++c) o += "_" + B.labels[c] + " #" + std::to_string(c + 1) + " \n";
            for (long r = 0; r < B.rows; ++r) {
                for (size_t c = 0; c < nl; ++c) {
                    const std::string v = escape(B.cells[(size_t)r * nl + c]);
                    if (v.size() < 10) o.append(10 - v.size(), ' ');
                    o += v;
                    o += ' ';
                }
                o += '\n';
            }
            o += " \n";
        } else {
            size_t w = 10;
            for (const auto &l : B.labels) w = std::max(w, l.size());
            for (size_t c = 0; c < nl; ++c) {
                o += "_" + B.labels[c];
                o.append(12 + w - B.labels[c].size(), ' ');  // std::setw(12 + maxWidth - w) << " "
                o += escape(B.rows ? B.cells[c] : std::string()) + "\n";
            }
            o += " \n";
        }
    }
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) RLN_FAIL(RLN_ERR_ARG, "MetaDataTable::write: cannot write to file: %s", path);
    const bool ok = std::fwrite(o.data(), 1, o.size(), f) == o.size();
    if (std::fclose(f) != 0 || !ok) RLN_FAIL(RLN_ERR_ARG, "write to %s failed", tmp.c_str());
    if (std::rename(tmp.c_str(), path) != 0) RLN_FAIL(RLN_ERR_ARG, "rename to %s failed: %s", path, strerror(errno));
    return RLN_OK;
}

// ------------------------------------------------------------------------------------------------ MRC C ABI
extern "C" int rln_mrc_read_info(const char *path, rln_mrc_info *info) {
    if (!path || !info) RLN_FAIL(RLN_ERR_ARG, "rln_mrc_read_info: null argument");
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) RLN_FAIL(RLN_ERR_ARG, "cannot open MRC file %s: %s", path, strerror(errno));
    MrcHeader h;
    return read_header(f.fd, path, h, *info);
}

extern "C" int rln_mrc_read(const char *path, long first, long count, float *out) {
    if (!path || (!out && count)) RLN_FAIL(RLN_ERR_ARG, "rln_mrc_read: null argument");
    Fd f;
    f.fd = open(path, O_RDONLY);
    if (f.fd < 0) RLN_FAIL(RLN_ERR_ARG, "cannot open MRC file %s: %s", path, strerror(errno));
    MrcHeader h;
    rln_mrc_info info;
    RLN_TRY(read_header(f.fd, path, h, info));
    return read_sections(f.fd, path, info, first, count, out);
}

extern "C" int rln_mrc_write(const char *path, const float *data, int nx, int ny, int nz, double pixel_size) {
    if (!path || !data || nx <= 0 || ny <= 0 || nz <= 0) RLN_FAIL(RLN_ERR_ARG, "rln_mrc_write: bad arguments");
    const size_t n = (size_t)nx * ny * nz;
    double mn = data[0], mx = data[0], sum = 0.;
    for (size_t k = 0; k < n; ++k) {
        mn = std::min(mn, (double)data[k]);
        mx = std::max(mx, (double)data[k]);
        sum += data[k];
    }
    const double avg = sum / (double)n;
    double ss = 0.;
    for (size_t k = 0; k < n; ++k) ss += ((double)data[k] - avg) * ((double)data[k] - avg);
    const double sd = n > 1 ? std::sqrt(ss / (double)(n - 1)) : 0.;  // computeStddev: over N - 1
    MrcHeader h;
    std::memset(&h, 0, sizeof(h));
    h.nx = nx;
    h.ny = ny;
    h.nz = nz;
    h.mode = 2;
    h.mx = nx;
    h.my = ny;
    h.mz = nz;
    const bool px = pixel_size > 0.;
    h.a = (float)(px ? pixel_size * nx : nx);
    h.b = (float)(px ? pixel_size * ny : ny);
    h.c = (float)(px ? pixel_size * nz : nz);
    h.alpha = h.beta = h.gamma = 90.f;
    h.mapc = 1;
    h.mapr = 2;
    h.maps = 3;
    h.amin = (float)mn;
    h.amax = (float)mx;
    h.amean = (float)avg;
    h.arms = (float)sd;
    std::memcpy(h.map, "MAP ", 4);
    h.machst[0] = 68;  // little-endian IEEE (this library's only target)
    h.machst[1] = 65;
    h.nlabl = 1;
    char label[80] = "relion_amd   ";
    const time_t now = time(nullptr);
    struct tm tmv;
    localtime_r(&now, &tmv);
    strftime(label + strlen(label), sizeof(label) - strlen(label), "%d-%b-%y  %R:%S", &tmv);
    std::memcpy(h.labels, label, sizeof(label));
    const std::string tmp = std::string(path) + ".tmp";
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) RLN_FAIL(RLN_ERR_ARG, "cannot write MRC file %s: %s", path, strerror(errno));
    bool ok = std::fwrite(&h, kMrcHeader, 1, f) == 1;
    ok = ok && std::fwrite(data, sizeof(float), n, f) == n;
    if (std::fclose(f) != 0 || !ok) RLN_FAIL(RLN_ERR_ARG, "write to %s failed", tmp.c_str());
    if (std::rename(tmp.c_str(), path) != 0) RLN_FAIL(RLN_ERR_ARG, "rename to %s failed: %s", path, strerror(errno));
    return RLN_OK;
}

// ------------------------------------------------------------------------------------------------ particle images
extern "C" int rln_star_read_images(const rln_star *s, int b, int c, long first, long count, const char *base_dir,
                                    int box, int n_threads, float *out) {
    if (!valid_block(s, b) || c < 0 || c >= (int)s->blocks[b].labels.size() || box <= 0 || (!out && count))
        RLN_FAIL(RLN_ERR_ARG, "rln_star_read_images: bad block, column, box or output");
    const auto &B = s->blocks[b];
    if (first < 0 || count < 0 || first + count > B.rows)
        RLN_FAIL(RLN_ERR_ARG, "rows %ld..%ld outside block %s of %ld rows", first, first + count - 1, B.name.c_str(), B.rows);
    const std::string base = (base_dir && *base_dir) ? std::string(base_dir) + "/" : std::string();
    // group the requested rows by stack file: (file) -> [(image index from 0, output slot)]
    std::map<std::string, std::vector<std::pair<long, long>>> groups;
    const size_t nl = B.labels.size();
    for (long r = 0; r < count; ++r) {
        const std::string &v = B.cells[(size_t)(first + r) * nl + c];
        const size_t at = v.find('@');
        long idx = 0;
        std::string file = v;
        if (at != std::string::npos) {
            char *end = nullptr;
            idx = std::strtol(v.c_str(), &end, 10) - 1;
            if (end != v.c_str() + at || idx < 0)
                RLN_FAIL(RLN_ERR_ARG, "image name '%s' (row %ld): the index before '@' is not a positive integer", v.c_str(),
                         first + r + 1);
            file = v.substr(at + 1);
        }
        groups[base + file].push_back({idx, r});
    }
    // headers once per stack, then runs of at most 32 consecutive images (consecutive in the stack and in the
    // output) spread over the threads, each opening its stack
    struct Stack {
        std::string path;
        rln_mrc_info info;
    };
    struct Run {
        int stack;
        long idx, slot, n;
    };
    std::vector<Stack> stacks;
    std::vector<Run> runs;
    for (auto &g : groups) {
        Stack st{g.first, rln_mrc_info{}};
        Fd f;
        f.fd = open(st.path.c_str(), O_RDONLY);
        if (f.fd < 0) RLN_FAIL(RLN_ERR_ARG, "cannot open image stack %s: %s", st.path.c_str(), strerror(errno));
        MrcHeader h;
        RLN_TRY(read_header(f.fd, st.path.c_str(), h, st.info));
        if (st.info.nx != box || st.info.ny != box)
            RLN_FAIL(RLN_ERR_ARG, "images of %s are %d x %d, not %d x %d", st.path.c_str(), st.info.nx, st.info.ny, box,
                     box);
        auto items = g.second;
        std::sort(items.begin(), items.end());
        for (const auto &it : items)
            if (it.first >= st.info.nz)
                RLN_FAIL(RLN_ERR_ARG, "readMRC: image number %ld exceeds stack size %d of image %s", it.first + 1,
                         st.info.nz, st.path.c_str());
        for (size_t k = 0; k < items.size();) {
            size_t e = k + 1;
            while (e < items.size() && e - k < 32 && items[e].first == items[e - 1].first + 1 &&
                   items[e].second == items[e - 1].second + 1)
                ++e;
            runs.push_back(Run{(int)stacks.size(), items[k].first, items[k].second, (long)(e - k)});
            k = e;
        }
        stacks.push_back(std::move(st));
    }
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
    nt = std::max(1, std::min<int>(nt, (int)runs.size()));
    std::vector<int> rc(runs.size(), RLN_OK);
    std::vector<std::string> msg(runs.size());
    const size_t px = (size_t)box * box;
#pragma omp parallel num_threads(nt)
    {
        int open_stack = -1;
        Fd f;
#pragma omp for schedule(dynamic, 4)
        for (long w = 0; w < (long)runs.size(); ++w) {
            const Run &r = runs[w];
            const Stack &st = stacks[r.stack];
            if (open_stack != r.stack) {
                if (f.fd >= 0) close(f.fd);
                f.fd = open(st.path.c_str(), O_RDONLY);
                open_stack = f.fd >= 0 ? r.stack : -1;
            }
            if (f.fd < 0) {
                rc[w] = RLN_ERR_ARG;
                msg[w] = "cannot open image stack " + st.path + ": " + strerror(errno);
                continue;
            }
            const int e = read_sections(f.fd, st.path.c_str(), st.info, r.idx, r.n, out + px * (size_t)r.slot);
            if (e != RLN_OK) {
                rc[w] = e;
                msg[w] = rln_last_error();
            }
        }
    }
    for (size_t w = 0; w < runs.size(); ++w)
        if (rc[w] != RLN_OK) RLN_FAIL(rc[w], "%s", msg[w].c_str());
    return RLN_OK;
}
