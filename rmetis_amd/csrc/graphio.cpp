// graphio.cpp -- METIS .graph files (the input format of the reference's
// README.md:208-218 and DEVELOPMENT_GUIDE.md:846-874): a header
// "nvtxs nedges [fmt] [ncon]" and one line per vertex holding, in order, its
// size (fmt digit 100), its ncon weights (digit 010) and its neighbours,
// 1-based, each followed by the edge weight (digit 001).  Lines starting
// with '%' are comments.  Every undirected edge is listed at both ends, so a
// file with nedges edges holds 2 * nedges entries.  Read into 0-based CSR.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "graphio.hpp"

namespace rk {

namespace {

struct Cursor {
    const char* p;
    const char* end;
    int64_t line = 1;

    // Skips blanks (not newlines).
    void blanks() {
        while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
    }
    // Moves to the start of the next line that is not a comment.
    void next_line() {
        while (p < end && *p != '\n') ++p;
        if (p < end) {
            ++p;
            ++line;
        }
    }
    void skip_comments() {
        for (;;) {
            const char* q = p;
            while (q < end && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
            if (q < end && *q == '%') {
                p = q;
                next_line();
                continue;
            }
            return;
        }
    }
    bool at_eol() {
        blanks();
        return p >= end || *p == '\n';
    }
    // Reads one integer of the current line; false at the end of the line.
    bool integer(int64_t* out) {
        blanks();
        if (p >= end || *p == '\n') return false;
        bool neg = false;
        if (*p == '-' || *p == '+') neg = *p++ == '-';
        if (p >= end || *p < '0' || *p > '9') throw Error(-2, "line " + std::to_string(line) + ": expected an integer");
        int64_t v = 0;
        while (p < end && *p >= '0' && *p <= '9') {
            v = v * 10 + (*p++ - '0');
            if (v > ((int64_t)1 << 40)) throw Error(-2, "line " + std::to_string(line) + ": integer out of range");
        }
        if (p < end && *p != ' ' && *p != '\t' && *p != '\r' && *p != '\n')
            throw Error(-2, "line " + std::to_string(line) + ": malformed number");
        *out = neg ? -v : v;
        return true;
    }
};

int32_t to_i32(int64_t v, const Cursor& c, const char* what) {
    if (v < INT32_MIN || v > INT32_MAX) throw Error(-2, "line " + std::to_string(c.line) + ": " + what + " out of range");
    return (int32_t)v;
}

}  // namespace

GraphFile read_graph_file(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw Error(-2, "cannot open " + path + ": " + std::strerror(errno));
    std::vector<char> buf;
    {
        std::fseek(f, 0, SEEK_END);
        const long sz = std::ftell(f);
        std::fseek(f, 0, SEEK_SET);
        if (sz < 0) {
            std::fclose(f);
            throw Error(-2, "cannot size " + path);
        }
        buf.resize((size_t)sz);
        const size_t got = sz > 0 ? std::fread(buf.data(), 1, (size_t)sz, f) : 0;
        std::fclose(f);
        if (got != (size_t)sz) throw Error(-2, "short read of " + path);
    }
    Cursor c{buf.data(), buf.data() + buf.size()};
    c.skip_comments();
    int64_t n = 0, m = 0, fmt = 0, ncon = 1;
    if (!c.integer(&n) || !c.integer(&m)) throw Error(-2, "the header needs nvtxs and nedges");
    if (c.integer(&fmt)) c.integer(&ncon);   // both optional
    if (!c.at_eol()) throw Error(-2, "line 1: extra fields in the header");
    if (n <= 0 || n >= INT32_MAX) throw Error(-2, "nvtxs out of range");
    if (m < 0 || 2 * m > INT32_MAX) throw Error(-2, "nedges out of range");
    if (fmt < 0 || fmt > 111 || (fmt % 10) > 1 || (fmt / 10 % 10) > 1 || fmt / 100 > 1)
        throw Error(-2, "fmt must be made of the digits 0 and 1 (vertex sizes, vertex weights, edge weights)");
    const bool has_vsize = fmt / 100 == 1, has_vwgt = fmt / 10 % 10 == 1, has_ewgt = fmt % 10 == 1;
    if (ncon < 1 || ncon > 1024) throw Error(-2, "ncon out of range");
    GraphFile g;
    g.n = (int32_t)n;
    g.ncon = (int32_t)ncon;
    g.xadj.reserve((size_t)n + 1);
    g.xadj.push_back(0);
    g.adjncy.reserve((size_t)(2 * m));
    if (has_ewgt) g.adjwgt.reserve((size_t)(2 * m));
    if (has_vwgt) g.vwgt.reserve((size_t)(n * ncon));
    if (has_vsize) g.vsize.reserve((size_t)n);
    g.has_vwgt = has_vwgt;
    g.has_adjwgt = has_ewgt;
    g.has_vsize = has_vsize;
    c.next_line();
    for (int64_t v = 0; v < n; ++v) {
        c.skip_comments();
        if (c.p >= c.end) throw Error(-2, "the file ends after " + std::to_string(v) + " of " + std::to_string(n) +
                                              " vertex lines");
        int64_t x = 0;
        if (has_vsize) {
            if (!c.integer(&x)) throw Error(-2, "line " + std::to_string(c.line) + ": missing vertex size");
            g.vsize.push_back(to_i32(x, c, "vertex size"));
        }
        if (has_vwgt)
            for (int64_t k = 0; k < ncon; ++k) {
                if (!c.integer(&x)) throw Error(-2, "line " + std::to_string(c.line) + ": missing vertex weight");
                g.vwgt.push_back(to_i32(x, c, "vertex weight"));
            }
        while (c.integer(&x)) {
            if (x < 1 || x > n)
                throw Error(-2, "line " + std::to_string(c.line) + ": neighbour " + std::to_string(x) +
                                    " outside 1.." + std::to_string(n));
            g.adjncy.push_back((int32_t)(x - 1));
            if (has_ewgt) {
                if (!c.integer(&x)) throw Error(-2, "line " + std::to_string(c.line) + ": missing edge weight");
                g.adjwgt.push_back(to_i32(x, c, "edge weight"));
            }
            if ((int64_t)g.adjncy.size() > 2 * m)
                throw Error(-2, "more adjacency entries than twice nedges (" + std::to_string(m) + ")");
        }
        g.xadj.push_back((int32_t)g.adjncy.size());
        c.next_line();
    }
    // only blank lines and comments may follow
    for (const char* was = nullptr; was != c.p;) {
        was = c.p;
        c.skip_comments();
        while (c.p < c.end && (*c.p == ' ' || *c.p == '\t' || *c.p == '\r' || *c.p == '\n')) {
            if (*c.p == '\n') ++c.line;
            ++c.p;
        }
    }
    if (c.p < c.end) throw Error(-2, "line " + std::to_string(c.line) + ": data after the last vertex line");
    if ((int64_t)g.adjncy.size() != 2 * m)
        throw Error(-2, "the vertex lines hold " + std::to_string(g.adjncy.size()) + " entries, not 2 * nedges = " +
                            std::to_string(2 * m));
    return g;
}

void write_graph_file(const std::string& path, int32_t n, int32_t ncon, const int32_t* xadj, const int32_t* adjncy,
                      const int32_t* vwgt, const int32_t* vsize, const int32_t* adjwgt) {
    if (n <= 0 || ncon <= 0 || !xadj || !adjncy) throw Error(-2, "nothing to write");
    if (xadj[n] % 2 != 0) throw Error(-2, "an undirected graph lists every edge twice (xadj[n] is odd)");
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw Error(-2, "cannot create " + path + ": " + std::strerror(errno));
    std::string out;
    out.reserve(1 << 20);
    const int fmt = (vsize ? 100 : 0) + (vwgt ? 10 : 0) + (adjwgt ? 1 : 0);
    out += std::to_string(n) + " " + std::to_string(xadj[n] / 2);
    if (fmt || ncon > 1) {
        char b[16];
        std::snprintf(b, sizeof(b), " %03d", fmt);
        out += b;
    }
    if (ncon > 1) out += " " + std::to_string(ncon);
    out += "\n";
    auto flush = [&]() {
        if (std::fwrite(out.data(), 1, out.size(), f) != out.size()) {
            std::fclose(f);
            throw Error(-4, "short write to " + path);
        }
        out.clear();
    };
    for (int32_t v = 0; v < n; ++v) {
        bool first = true;
        auto field = [&](int64_t x) {
            if (!first) out += ' ';
            out += std::to_string(x);
            first = false;
        };
        if (vsize) field(vsize[v]);
        if (vwgt)
            for (int32_t k = 0; k < ncon; ++k) field(vwgt[(int64_t)v * ncon + k]);
        for (int32_t j = xadj[v]; j < xadj[v + 1]; ++j) {
            field((int64_t)adjncy[j] + 1);
            if (adjwgt) field(adjwgt[j]);
        }
        out += '\n';
        if (out.size() > (1u << 20)) flush();
    }
    flush();
    if (std::fclose(f) != 0) throw Error(-4, "cannot close " + path);
}

}  // namespace rk
