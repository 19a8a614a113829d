// The library's CPU-only code (csrc/host.cpp, csrc/plan.cpp, csrc/sigproc.cpp)
// built on its own with AddressSanitizer and UndefinedBehaviorSanitizer and
// driven through its C ABI: ddplan/JSON parsing (valid and broken inputs),
// ndump, clustering, kadane and the SIGPROC reader.  The three entry points
// host.cpp reaches into the device library for are stubbed to fail.
// argv[1], argv[2]: the reference's example ddplan and JSON config.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "txgpu.h"

namespace txgpu {
int set_error(int code, const std::string &) { return code; }
}
extern "C" int tx_get_info(const tx_plan *, tx_info *) { return TX_EINVAL; }
extern "C" int tx_get_tables(const tx_plan *, int *, int *, double *, double *, double *, double *)
{
    return TX_EINVAL;
}
extern "C" int tx_cand_spectra(tx_plan *, int, const int *, const int *, const int *, float *) { return TX_EINVAL; }

static std::string slurp(const char *path)
{
    FILE *f = std::fopen(path, "rb");
    if (!f) return "";
    std::string s;
    char buf[4096];
    size_t k;
    while ((k = std::fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    std::fclose(f);
    return s;
}

#define CHECK(c)                                                                                   \
    do {                                                                                           \
        if (!(c)) {                                                                                \
            std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c);                     \
            return 1;                                                                              \
        }                                                                                          \
    } while (0)

int main(int argc, char **argv)
{
    if (argc < 3) return 2;
    tx_ddrow def{};
    def.td = def.fd = 1;
    def.ddm = 1;
    def.ndm = 100;
    def.minw = 1e-4f;
    def.maxw = 0.1f;
    def.snrloss = 0.1f;
    def.thre = 7;
    std::vector<tx_ddrow> rows(64);
    int n = 0;
    // the reference's own examples
    CHECK(tx_ddplan_parse(slurp(argv[1]).c_str(), 0, &def, rows.data(), 64, &n) == TX_OK && n >= 1);
    CHECK(tx_ddplan_parse(slurp(argv[2]).c_str(), 1, &def, rows.data(), 64, &n) == TX_OK && n >= 1);
    // too small a row buffer, broken text and JSON, long rfi/zap lists
    CHECK(tx_ddplan_parse(slurp(argv[1]).c_str(), 0, &def, rows.data(), 1, &n) != TX_OK || n <= 1);
    const char *bad[] = {"1 1 0 1\n", "a b c d e f g\n", "{", "{\"ddplan\": 3}", "[1,2", "{\"ddplan\": [{\"td\": }]}",
                         "\"\\", "{\"ddplan\": [[[[[[[[[[[[[[[[[[[[]]]]]]]]]]]]]]]]]]]]}"};
    for (const char *b : bad) tx_ddplan_parse(b, b[0] == '{' || b[0] == '[' || b[0] == '"', &def, rows.data(), 64, &n);
    std::string many = "#\n";
    for (int i = 0; i < 200; i++) many += "1 1 0 1 100 0.1 0.1 zap 1000 1100 zdot kadaneF 4 8 mask 2 2\n";
    tx_ddplan_parse(many.c_str(), 0, &def, rows.data(), 64, &n);
    CHECK(tx_ddplan_parse(nullptr, 0, &def, rows.data(), 64, &n) == TX_EINVAL);

    // ndump
    const int tds[] = {1, 2, 4, 3};
    CHECK(tx_ndump(2.0f, 76.56e-6, tds, 4) > 0);
    CHECK(tx_ndump(2.0f, 76.56e-6, tds, 0) < 0 || tx_ndump(2.0f, 76.56e-6, tds, 0) > 0);
    CHECK(tx_ndump(2.0f, -1.0, tds, 4) < 0);

    // clustering: random points in row-major order, several caps
    std::mt19937 rng(7);
    for (int trial = 0; trial < 20; trial++) {
        const int np = trial == 0 ? 0 : (int)(rng() % 3000);
        std::vector<tx_point> pts;
        for (int idm = 0; idm < 200 && (int)pts.size() < np; idm++)
            for (int t = 0; t < 5000 && (int)pts.size() < np; t++)
                if (rng() % 97 == 0) pts.push_back(tx_point{idm, t, 7.0f + (float)(rng() % 1000) / 100.f, 1 + (int)(rng() % 64)});
        tx_cluster_desc d{};
        d.dms = 0;
        d.ddm = 0.5;
        d.fmax = 1712;
        d.fmin = 856;
        d.tsamp = 76.56e-6;
        d.radius = 1e-3 * (1 + trial % 4);
        d.maxncand = trial % 3 == 0 ? 5 : 100;
        d.minpts = trial % 5;
        d.drop = trial % 2;
        d.maxwn = 64;
        const size_t cap = trial % 4 == 0 ? 3 : 4096;
        std::vector<tx_cand> out(cap);
        std::vector<int32_t> labels(pts.size() + 1);
        size_t nc = 0;
        const int rc = tx_cluster(pts.data(), pts.size(), &d, out.data(), cap, &nc, labels.data());
        CHECK(rc == TX_OK || rc == TX_EOVERFLOW);
    }

    // kadane
    for (int trial = 0; trial < 50; trial++) {
        const long len = trial == 0 ? 0 : (long)(rng() % 300);
        std::vector<float> a((size_t)len + 1);
        for (long i = 0; i < len; i++) a[i] = (float)((int)(rng() % 200) - 100) / 10.f;
        long s = -1, e = -1;
        tx_kadane(a.data(), len, &s, &e);
    }

    // SIGPROC: write, read in odd pieces, skip, refuse a truncated copy
    const char *path = "host_sanitize.fil";
    tx_fil_header h{};
    std::strcpy(h.source_name, "J0000+00");
    h.nbits = 8;
    h.nchans = 48;
    h.nifs = 1;
    h.nbeams = 1;
    h.tsamp = 1e-4;
    h.fch1 = 1500;
    h.foff = -1;
    h.tstart = 59000;
    CHECK(tx_fil_write_header(path, &h) == TX_OK);
    {
        FILE *f = std::fopen(path, "ab");
        std::vector<unsigned char> d(48 * 1001 + 7);
        for (auto &x : d) x = (unsigned char)rng();
        std::fwrite(d.data(), 1, d.size(), f);
        std::fclose(f);
    }
    const char *paths[] = {path, path};
    tx_fil *r = nullptr;
    CHECK(tx_fil_open(paths, 2, 0, &r, &h) == TX_OK && h.nsamples == 2002);
    std::vector<unsigned char> buf(48 * 700);
    long got = 0, k;
    while ((k = tx_fil_read(r, buf.data(), 1 + (long)(rng() % 700))) > 0) got += k;
    CHECK(got == 2002);
    CHECK(tx_fil_skip(r, 5, 9) == TX_OK);
    got = 0;
    while ((k = tx_fil_read(r, buf.data(), 333)) > 0) got += k;
    CHECK(got == 2002 - 14);
    CHECK(tx_fil_skip(r, 2000, 5) != TX_OK);
    std::vector<double> fr(48);
    CHECK(tx_fil_frequencies(r, fr.data()) == TX_OK && fr[47] == 1453);
    tx_fil_close(r);
    CHECK(tx_fil_open(paths, 2, 1, &r, &h) != TX_OK); // the second file does not continue the first
    {
        std::string whole = slurp(path);
        for (size_t cut : {size_t(0), size_t(3), size_t(16), whole.size() / 8, size_t(200)}) {
            FILE *f = std::fopen("host_sanitize_cut.fil", "wb");
            std::fwrite(whole.data(), 1, std::min(cut, whole.size()), f);
            std::fclose(f);
            tx_fil_header hc{};
            tx_fil_read_header("host_sanitize_cut.fil", &hc);
        }
    }
    std::remove(path);
    std::remove("host_sanitize_cut.fil");
    std::printf("host_sanitize ok\n");
    return 0;
}
