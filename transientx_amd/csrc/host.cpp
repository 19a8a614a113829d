// host.cpp -- host side of the single-pulse search around the kernels:
//   * ddplan text / JSON config -> search rows, ndump (td_lcm)
//   * the cluster of thresholded points (Cluster::run + clustering())
//   * fl/fh of a candidate (kadane over the sub-band spectrum)
//   * the .cands line of CandPlot::plot
// Everything here is plain C++17 on the CPU; the only device work it starts
// is through the public C ABI (tx_cand_spectra).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/txgpu.h"
#include "plan.h"

namespace txgpu {
int set_error(int code, const std::string &msg);
}
using txgpu::set_error;

// ------------------------------------------------------------------ ddplan
namespace {

// Front-end RFI items ("mask a b", "kadaneF a b", "kadaneT a b", "zero",
// "zdot") and zap ranges ("fl fh") as ';'-separated text in a fixed field.
bool append_item(char *dst, size_t cap, const std::string &item)
{
    std::string cur(dst);
    if (!cur.empty()) cur += ";";
    cur += item;
    if (cur.size() + 1 > cap) return false;
    std::memcpy(dst, cur.c_str(), cur.size() + 1);
    return true;
}

// std::stoi / stod / stof semantics: leading number, the rest ignored,
// nothing numeric is an error.
bool num_d(const std::string &s, double &v)
{
    const char *b = s.c_str();
    char *e = nullptr;
    v = std::strtod(b, &e);
    return e != b;
}
bool num_l(const std::string &s, long &v)
{
    const char *b = s.c_str();
    char *e = nullptr;
    v = std::strtol(b, &e, 10);
    return e != b;
}

// RFI options of one ddplan line or of the command line
// (src/singlepulse/singlepulse.cpp:352-373 and :428-446): unknown tokens are
// skipped, a named item takes its arguments with it.
int parse_rfi_tokens(const std::vector<std::string> &t, size_t from, tx_ddrow &row)
{
    for (size_t k = from; k < t.size(); k++) {
        const std::string &o = t[k];
        if (o == "mask" || o == "kadaneF" || o == "kadaneT") {
            if (k + 2 >= t.size()) return set_error(TX_EINVAL, "rfi item '" + o + "' needs two values");
            if (!append_item(row.rfi, sizeof(row.rfi), o + " " + t[k + 1] + " " + t[k + 2]))
                return set_error(TX_ERANGE, "rfi list too long");
            k += 2;
        } else if (o == "zap") {
            double a, b;
            if (k + 2 >= t.size() || !num_d(t[k + 1], a) || !num_d(t[k + 2], b))
                return set_error(TX_EINVAL, "zap needs two frequencies");
            char tmp[96];
            std::snprintf(tmp, sizeof(tmp), "%.17g %.17g", a, b);
            if (!append_item(row.zap, sizeof(row.zap), tmp)) return set_error(TX_ERANGE, "zap list too long");
            k += 2;
        } else if (o == "zero" || o == "zdot") {
            if (!append_item(row.rfi, sizeof(row.rfi), o)) return set_error(TX_ERANGE, "rfi list too long");
        }
    }
    return TX_OK;
}

// ---- a small JSON reader (objects, arrays, strings, numbers, literals)
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0;
    std::string str;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;
    const JVal *get(const char *key) const
    {
        if (kind != OBJ) return nullptr;
        for (auto &kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
};

struct JParser {
    const char *p, *end;
    std::string err;
    void ws()
    {
        while (p < end && std::isspace((unsigned char)*p)) p++;
    }
    bool lit(const char *s)
    {
        size_t n = std::strlen(s);
        if ((size_t)(end - p) >= n && std::strncmp(p, s, n) == 0) {
            p += n;
            return true;
        }
        return false;
    }
    bool str(std::string &out)
    {
        if (p >= end || *p != '"') return false;
        p++;
        out.clear();
        while (p < end && *p != '"') {
            char c = *p++;
            if (c == '\\' && p < end) {
                char e = *p++;
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (end - p < 4) return false;
                    unsigned v = (unsigned)std::strtoul(std::string(p, 4).c_str(), nullptr, 16);
                    p += 4;
                    if (v < 0x80) out += (char)v;
                    else if (v < 0x800) {
                        out += (char)(0xC0 | (v >> 6));
                        out += (char)(0x80 | (v & 0x3F));
                    } else {
                        out += (char)(0xE0 | (v >> 12));
                        out += (char)(0x80 | ((v >> 6) & 0x3F));
                        out += (char)(0x80 | (v & 0x3F));
                    }
                    break;
                }
                default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (p >= end) return false;
        p++;
        return true;
    }
    bool value(JVal &v, int depth)
    {
        if (depth > 64) return false;
        ws();
        if (p >= end) return false;
        if (*p == '{') {
            p++;
            v.kind = JVal::OBJ;
            ws();
            if (p < end && *p == '}') {
                p++;
                return true;
            }
            for (;;) {
                ws();
                std::string k;
                if (!str(k)) return false;
                ws();
                if (p >= end || *p != ':') return false;
                p++;
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.obj.emplace_back(std::move(k), std::move(x));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == '}') {
                    p++;
                    return true;
                }
                return false;
            }
        }
        if (*p == '[') {
            p++;
            v.kind = JVal::ARR;
            ws();
            if (p < end && *p == ']') {
                p++;
                return true;
            }
            for (;;) {
                JVal x;
                if (!value(x, depth + 1)) return false;
                v.arr.push_back(std::move(x));
                ws();
                if (p < end && *p == ',') {
                    p++;
                    continue;
                }
                if (p < end && *p == ']') {
                    p++;
                    return true;
                }
                return false;
            }
        }
        if (*p == '"') {
            v.kind = JVal::STR;
            return str(v.str);
        }
        if (lit("true")) {
            v.kind = JVal::BOOL;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.kind = JVal::BOOL;
            v.b = false;
            return true;
        }
        if (lit("null")) {
            v.kind = JVal::NUL;
            return true;
        }
        char *e = nullptr;
        v.num = std::strtod(p, &e);
        if (e == p) return false;
        v.kind = JVal::NUM;
        p = e;
        return true;
    }
};

// nlohmann::json's implicit conversions as SinglePulse's fields use them
// (src/singlepulse/singlepulse.cpp:474-509): a missing key is an error.
bool jnum(const JVal *o, const char *sec, const char *key, double &v, std::string &err)
{
    const JVal *s = o ? o->get(sec) : nullptr;
    const JVal *x = s ? s->get(key) : nullptr;
    if (!x || (x->kind != JVal::NUM && x->kind != JVal::BOOL)) {
        err = std::string("ddplan entry needs a number at ") + sec + "." + key;
        return false;
    }
    v = x->kind == JVal::BOOL ? (x->b ? 1. : 0.) : x->num;
    return true;
}

} // namespace

extern "C" int tx_ddplan_parse(const char *text, int is_json, const tx_ddrow *defaults,
                               tx_ddrow *rows, int cap, int *n)
{
    if (!text || !defaults || !n || (cap > 0 && !rows)) return set_error(TX_EINVAL, "null argument");
    *n = 0;
    // one SinglePulse is filled and copied per row, so rfi/zap items
    // accumulate from row to row (src/singlepulse/singlepulse.cpp:345,448 and :455-543)
    tx_ddrow sp = *defaults;
    std::vector<tx_ddrow> out;
    if (!is_json) {
        // src/singlepulse/singlepulse.cpp:403-452
        const char *p = text;
        int id = 0;
        while (*p) {
            const char *e = std::strchr(p, '\n');
            std::string line = e ? std::string(p, e - p) : std::string(p);
            p = e ? e + 1 : p + line.size();
            // boost::trim, then lines not starting with a digit are comments
            size_t b = 0, t = line.size();
            while (b < t && std::isspace((unsigned char)line[b])) b++;
            while (t > b && std::isspace((unsigned char)line[t - 1])) t--;
            line = line.substr(b, t - b);
            if (line.empty() || !std::isdigit((unsigned char)line[0])) continue;
            std::vector<std::string> tok;
            size_t i = 0;
            while (i < line.size()) {
                while (i < line.size() && (line[i] == ' ' || line[i] == '\t')) i++;
                size_t j = i;
                while (j < line.size() && line[j] != ' ' && line[j] != '\t') j++;
                if (j > i) tok.push_back(line.substr(i, j - i));
                i = j;
            }
            if (tok.size() < 7) return set_error(TX_EINVAL, "ddplan line needs td fd dms ddm ndm snrloss maxw: " + line);
            long td, fd, ndm;
            double dms, ddm, snrloss, maxw;
            if (!num_l(tok[0], td) || !num_l(tok[1], fd) || !num_d(tok[2], dms) || !num_d(tok[3], ddm) ||
                !num_l(tok[4], ndm) || !num_d(tok[5], snrloss) || !num_d(tok[6], maxw))
                return set_error(TX_EINVAL, "ddplan line is not numeric: " + line);
            sp.td = (int)td;
            sp.fd = (int)fd;
            sp.dms = dms;
            sp.ddm = ddm;
            sp.ndm = ndm;
            sp.snrloss = (float)snrloss;
            sp.maxw = (float)maxw;
            int rc = parse_rfi_tokens(tok, 7, sp);
            if (rc) return rc;
            sp.id = ++id;
            out.push_back(sp);
        }
        if (out.empty()) {
            // no ddplan rows: one row from the command line (singlepulse.cpp:453-457)
            sp.id = 1;
            out.push_back(sp);
        }
    } else {
        // src/singlepulse/singlepulse.cpp:460-545
        JParser jp{text, text + std::strlen(text), {}};
        JVal root;
        if (!jp.value(root, 0)) return set_error(TX_EINVAL, "config is not valid JSON");
        const JVal *plan = root.get("ddplan");
        if (!plan || plan->kind != JVal::ARR) return set_error(TX_EINVAL, "config has no ddplan array");
        int id = 0;
        for (const JVal &c : plan->arr) {
            std::string err;
            double v[24];
            struct K {
                const char *sec, *key;
            };
            static const K keys[] = {
                {"downsample", "td"},    {"downsample", "fd"},        {"baseline", "width"},
                {"subdedispersion", "dms"}, {"subdedispersion", "ddm"}, {"subdedispersion", "ndm"},
                {"subdedispersion", "overlap"}, {"rfi", "bandlimit"},  {"rfi", "bandlimitKT"},
                {"rfi", "widthlimit"},   {"rfi", "threMask"},          {"rfi", "threKadaneF"},
                {"rfi", "threKadaneT"},  {"boxcar", "minw"},           {"boxcar", "snrloss"},
                {"boxcar", "maxw"},      {"boxcar", "iqr"},            {"clustering", "thre"},
                {"clustering", "radius"}, {"clustering", "neighbors"}, {"clustering", "maxncand"},
                {"clustering", "minpts"}, {"clustering", "drop"}};
            const int nk = (int)(sizeof(keys) / sizeof(keys[0]));
            for (int k = 0; k < nk; k++)
                if (!jnum(&c, keys[k].sec, keys[k].key, v[k], err)) return set_error(TX_EINVAL, err);
            sp.td = (int)v[0];
            sp.fd = (int)v[1];
            sp.bswidth = (float)v[2];
            sp.dms = v[3];
            sp.ddm = v[4];
            sp.ndm = (long)v[5];
            sp.overlap = v[6];
            sp.bandlimit = v[7];
            sp.bandlimitKT = v[8];
            sp.widthlimit = v[9];
            sp.threMask = (float)v[10];
            sp.threKadaneF = (float)v[11];
            sp.threKadaneT = (float)v[12];
            sp.minw = (float)v[13];
            sp.snrloss = (float)v[14];
            sp.maxw = (float)v[15];
            sp.iqr = v[16] != 0.;
            sp.thre = (float)v[17];
            sp.radius = v[18] / 1000.;
            sp.neighbors = (int)v[19];
            sp.maxncand = (int)v[20];
            sp.minpts = (int)v[21];
            sp.drop = v[22] != 0.;
            // optional additions: sub-bands and fine DMs per coarse DM
            const JVal *dd = c.get("subdedispersion");
            const JVal *ns = dd ? dd->get("nsub") : nullptr;
            if (ns && ns->kind == JVal::NUM) sp.nsub = (int)ns->num;
            const JVal *gr = dd ? dd->get("group") : nullptr;
            if (gr && gr->kind == JVal::NUM) sp.group = (int)gr->num;
            const JVal *rfi = c.get("rfi");
            const JVal *zl = rfi ? rfi->get("zaplist") : nullptr;
            if (zl && zl->kind == JVal::ARR)
                for (const JVal &z : zl->arr) {
                    if (z.kind != JVal::ARR || z.arr.empty() || z.arr.front().kind != JVal::NUM ||
                        z.arr.back().kind != JVal::NUM)
                        return set_error(TX_EINVAL, "zaplist items are [fl, fh]");
                    char tmp[96];
                    std::snprintf(tmp, sizeof(tmp), "%.17g %.17g", z.arr.front().num, z.arr.back().num);
                    if (!append_item(sp.zap, sizeof(sp.zap), tmp)) return set_error(TX_ERANGE, "zap list too long");
                }
            const JVal *rl = rfi ? rfi->get("rfilist") : nullptr;
            if (rl && rl->kind == JVal::ARR)
                for (const JVal &r : rl->arr) {
                    if (r.kind != JVal::ARR) return set_error(TX_EINVAL, "rfilist items are arrays");
                    std::vector<std::string> t;
                    for (const JVal &x : r.arr) {
                        if (x.kind != JVal::STR) return set_error(TX_EINVAL, "rfilist entries are strings");
                        t.push_back(x.str);
                    }
                    for (size_t k = 0; k < t.size(); k++) {
                        const std::string &o = t[k];
                        if (o == "mask" || o == "kadaneF" || o == "kadaneT") {
                            if (k + 2 >= t.size()) return set_error(TX_EINVAL, "rfi item '" + o + "' needs two values");
                            if (!append_item(sp.rfi, sizeof(sp.rfi), o + " " + t[k + 1] + " " + t[k + 2]))
                                return set_error(TX_ERANGE, "rfi list too long");
                            k += 2;
                        } else if (o == "zero" || o == "zdot") {
                            if (!append_item(sp.rfi, sizeof(sp.rfi), o))
                                return set_error(TX_ERANGE, "rfi list too long");
                        }
                    }
                }
            sp.id = ++id;
            out.push_back(sp);
        }
    }
    *n = (int)out.size();
    if ((int)out.size() > cap) return set_error(TX_EOVERFLOW, "more ddplan rows than cap");
    std::copy(out.begin(), out.end(), rows);
    return TX_OK;
}

// src/singlepulse/dedisperse_fil.cpp:224-244
extern "C" long tx_ndump(float seglen, double tsamp, const int *tds, int n)
{
    if (!tds || n <= 0 || tsamp <= 0.) return set_error(TX_EINVAL, "bad ndump arguments");
    long lcm = 1;
    for (int i = 0; i < n; i++) {
        if (tds[i] <= 0) return set_error(TX_EINVAL, "td must be positive");
        lcm = lcm / std::gcd(lcm, (long)tds[i]) * tds[i];
    }
    return (long)(int)(seglen / tsamp) / lcm * lcm;
}

// ------------------------------------------------------------------ cluster
// Cluster<T>::run (src/singlepulse/cluster.cpp:63-161) over the thresholded
// points in row-major order, with clustering() of include/clustering.h:26-100:
//   * seeds are taken by the largest S among unclustered points (strict >,
//     so the earliest wins); the seed's index is held in a float
//     (clustering.h:44), so above 2^24 points the seed position is that of
//     the index rounded to float;
//   * the seed's neighbours within radius (inclusive, distance^2 <= r^2) of
//     any state join the cluster, then the region grows through unclustered
//     neighbours only;
//   * at most maxncand clusters;
//   * per cluster the point of largest S (strict >, from 0) is the candidate;
//     a cluster is dropped when its size < minpts or (drop and its width is
//     maxwn), and the drop erases candcluster[k] while k runs on, so later
//     clusters are judged by the size of the cluster one position further on
//     (cluster.cpp:145-158); past the shortened list the size reads as 0.
// Coordinates are (dmdelay(dm, fmax, fmin)/2, isamp*tsamp) in seconds.
extern "C" int tx_cluster(const tx_point *pts, size_t n, const tx_cluster_desc *d, tx_cand *out,
                          size_t cap, size_t *ncand, int32_t *labels)
{
    if ((!pts && n) || !d || !ncand || (cap && !out)) return set_error(TX_EINVAL, "null argument");
    *ncand = 0;
    if (d->tsamp <= 0.) return set_error(TX_EINVAL, "tsamp must be positive");
    double radius = d->radius;
    if (radius < 1.5 * d->tsamp) radius = 1.5 * d->tsamp;
    const double r2 = radius * radius;
    if (labels)
        for (size_t i = 0; i < n; i++) labels[i] = 0;
    if (n == 0) return TX_OK;

    std::vector<double> x(n), y(n);
    for (size_t i = 0; i < n; i++) {
        const double dm = d->dms + pts[i].idm * d->ddm;
        x[i] = txgpu::dmdelay(dm, d->fmax, d->fmin) * 0.5;
        y[i] = pts[i].isamp * d->tsamp;
    }
    // uniform grid with cell = radius: neighbours lie in the 3x3 cells around
    auto cell = [&](double v) { return (long long)std::floor(v / radius); };
    auto key = [](long long cx, long long cy) {
        return ((unsigned long long)cx * 0x9E3779B97F4A7C15ull) ^ (unsigned long long)cy;
    };
    std::unordered_map<unsigned long long, std::vector<long>> grid;
    grid.reserve(n);
    std::vector<long long> cx(n), cy(n);
    for (size_t i = 0; i < n; i++) {
        cx[i] = cell(x[i]);
        cy[i] = cell(y[i]);
        grid[key(cx[i], cy[i])].push_back((long)i);
    }
    std::vector<int> label(n, 0);
    std::vector<char> done(n, 0);
    std::vector<long> front, next;
    std::vector<char> mark(n, 0);
    auto neighbours = [&](double px, double py, long long qx, long long qy, std::vector<long> &acc) {
        for (long long a = -1; a <= 1; a++)
            for (long long b = -1; b <= 1; b++) {
                auto it = grid.find(key(qx + a, qy + b));
                if (it == grid.end()) continue;
                for (long q : it->second) {
                    if (cx[q] != qx + a || cy[q] != qy + b) continue;
                    const double dx = x[q] - px, dy = y[q] - py;
                    if (dx * dx + dy * dy <= r2 && !mark[q]) {
                        mark[q] = 1;
                        acc.push_back(q);
                    }
                }
            }
    };
    size_t remaining = n;
    int icluster = 0;
    while ((icluster++ < d->maxncand) && remaining > 0) {
        float maxsnr = -1.f;
        long seed = -1;
        for (size_t i = 0; i < n; i++)
            if (!done[i] && pts[i].S > maxsnr) {
                maxsnr = pts[i].S;
                seed = (long)i;
            }
        if (seed < 0) break; // no point above -1: the reference would index -1
        const long s = (long)(float)seed;
        front.clear();
        neighbours(x[s], y[s], cx[s], cy[s], front);
        for (long q : front) mark[q] = 0;
        while (!front.empty()) {
            next.clear();
            for (long q : front) {
                label[q] = icluster;
                if (!done[q]) {
                    done[q] = 1;
                    remaining--;
                }
                mark[q] = 1; // in this level: not re-collected below
            }
            for (long q : front) neighbours(x[q], y[q], cx[q], cy[q], next);
            for (long q : front) mark[q] = 0;
            for (long q : next) mark[q] = 0;
            front.clear();
            for (long q : next)
                if (!done[q]) front.push_back(q);
        }
    }
    const int ncluster = icluster - 1;
    std::vector<long> best(ncluster > 0 ? ncluster : 0, -1), size(ncluster > 0 ? ncluster : 0, 0);
    std::vector<float> bestS(best.size(), 0.f);
    for (size_t i = 0; i < n; i++) {
        const int c = label[i];
        if (c == 0) continue;
        size[c - 1]++;
        if (pts[i].S > bestS[c - 1]) {
            bestS[c - 1] = pts[i].S;
            best[c - 1] = (long)i;
        }
    }
    if (labels)
        for (size_t i = 0; i < n; i++) labels[i] = label[i];
    std::vector<long> live = size; // candcluster sizes as the erase loop sees them
    std::vector<tx_cand> cands;
    for (int k = 0; k < ncluster; k++) {
        const long bi = best[k];
        bool keep;
        if (bi < 0) keep = false;
        else if (d->drop && pts[bi].wn == d->maxwn) keep = false;
        else {
            const unsigned long sz = k < (long)live.size() ? (unsigned long)live[k] : 0ul;
            keep = sz >= (unsigned long)(long)d->minpts;
        }
        if (keep) {
            tx_cand c;
            c.idm = pts[bi].idm;
            c.isamp = pts[bi].isamp;
            c.wn = pts[bi].wn;
            c.S = pts[bi].S;
            c.cluster = k + 1;
            c.npts = (int32_t)size[k];
            cands.push_back(c);
        } else if (k < (long)live.size()) {
            live.erase(live.begin() + k);
        }
    }
    *ncand = cands.size();
    if (cands.size() > cap) return set_error(TX_EOVERFLOW, "more candidates than cap");
    std::copy(cands.begin(), cands.end(), out);
    return TX_OK;
}

// ------------------------------------------------------------ fl/fh, .cands
// kadane<float> as src/singlepulse/candplot.cpp:229 calls it (XLibs; declared:
// the running sum restarts after it drops below zero, strict '>' keeps the
// earliest maximum).
extern "C" void tx_kadane(const float *a, long n, long *start, long *end)
{
    float best = -INFINITY, cur = 0.f;
    long s = 0, bs = 0, be = 0;
    for (long i = 0; i < n; i++) {
        cur += a[i];
        if (cur > best) {
            best = cur;
            bs = s;
            be = i;
        }
        if (cur < 0.f) {
            cur = 0.f;
            s = i + 1;
        }
    }
    *start = bs;
    *end = be;
}

// src/singlepulse/candplot.cpp:141-231: fl = vfnorm[fstart], fh = vfnorm[fend],
// vfnorm the sub-band frequencies re-spaced linearly from first to last.
extern "C" int tx_cand_flfh(tx_plan *p, int ncand, const int *idm, const int *isamp, const int *wn,
                            float *fl, float *fh)
{
    if (!p || ncand < 0 || (ncand && (!idm || !isamp || !wn || !fl || !fh)))
        return set_error(TX_EINVAL, "null argument");
    if (ncand == 0) return TX_OK;
    tx_info info;
    int rc = tx_get_info(p, &info);
    if (rc) return rc;
    const int nsub = info.nsub;
    std::vector<double> fsub(nsub);
    rc = tx_get_tables(p, nullptr, nullptr, nullptr, nullptr, fsub.data(), nullptr);
    if (rc) return rc;
    std::vector<float> spec((size_t)ncand * nsub);
    rc = tx_cand_spectra(p, ncand, idm, isamp, wn, spec.data());
    if (rc) return rc;
    std::vector<float> vfnorm(nsub);
    for (int j = 0; j < nsub; j++)
        vfnorm[j] = (float)(fsub.front() + j / (nsub - 1.) * (fsub.back() - fsub.front()));
    for (int c = 0; c < ncand; c++) {
        long s = 0, e = 0;
        tx_kadane(spec.data() + (size_t)c * nsub, nsub, &s, &e);
        fl[c] = vfnorm[s];
        fh[c] = vfnorm[e];
    }
    return TX_OK;
}

// CandPlot::plot's .cands lines (src/singlepulse/candplot.cpp:32-48,420-433),
// candidates in stable order of descending S (include/candplot.h:27-33).
extern "C" int tx_cands_format(tx_plan *p, const tx_cand *cands, size_t n, const tx_cands_ctx *ctx,
                               char *buf, size_t cap, size_t *len)
{
    if (!p || !ctx || !len || (n && !cands) || (cap && !buf)) return set_error(TX_EINVAL, "null argument");
    *len = 0;
    tx_info info;
    int rc = tx_get_info(p, &info);
    if (rc) return rc;
    const double dms = info.dms, ddm = info.ddm, tsamp = info.tsamp;
    std::vector<size_t> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cands[a].S > cands[b].S; });
    std::vector<int> vi(n), vs(n), vw(n);
    for (size_t k = 0; k < n; k++) {
        vi[k] = cands[k].idm;
        vs[k] = cands[k].isamp;
        vw[k] = cands[k].wn;
    }
    std::vector<float> fl(n), fh(n);
    rc = tx_cand_flfh(p, (int)n, vi.data(), vs.data(), vw.data(), fl.data(), fh.data());
    if (rc) return rc;
    const char *beam = ctx->beam ? ctx->beam : "";
    const char *root = ctx->rootname ? ctx->rootname : "";
    const char *fname = ctx->fname ? ctx->fname : "";
    std::string text;
    long num = ctx->num_cand ? *ctx->num_cand : 0;
    for (size_t iidx = 0; iidx < n; iidx++) {
        const size_t k = idx[iidx];
        const tx_cand &c = cands[k];
        const double mjd = ctx->tstart + (-info.offset + info.counter - info.ndump + (long)c.isamp) * tsamp / 86400.;
        const double bmjd = ctx->tstart + (-info.offset + info.counter - info.ndump) * tsamp / 86400.;
        char png[4096];
        std::snprintf(png, sizeof(png), "%s_%.10f_%s_%02ld_%02d.png", root, bmjd, beam, (long)(iidx + 1),
                      ctx->plan_id);
        char line[8192];
        std::snprintf(line, sizeof(line), "%s\t%ld\t%.15f\t%.2f\t%.2f\t%.1f\t%.0f\t%.0f\t%s\t%02d\t%s\n", beam,
                      ++num, mjd, dms + c.idm * ddm, c.wn * tsamp * 1000, (double)c.S, (double)fl[k],
                      (double)fh[k], png, ctx->plan_id, fname);
        text += line;
    }
    if (ctx->num_cand) *ctx->num_cand = num;
    *len = text.size();
    if (text.size() + 1 > cap) return set_error(TX_EOVERFLOW, "buffer too small for the .cands lines");
    std::memcpy(buf, text.c_str(), text.size() + 1);
    return TX_OK;
}
