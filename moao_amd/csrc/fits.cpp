// FITS primary images in the layout the reference's tools read and write (cfitsio through
// common/io/fits.cpp:66-148 and common/tomo_struct/intersample.cpp:210-300): a primary-HDU image,
// NAXIS1 the fast axis, data big-endian in memory order, no scaling keywords.  Only what those
// files need: BITPIX -64/-32 (IEEE) and 8/16/32/64 (integers), read as doubles (an integer image
// such as idx.fits holds values far below 2^53, so the conversion is exact); written as doubles
// (writeFits writes every matrix with TDOUBLE).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace moao {

int fits_read(const std::string& path, long* nx, long* ny, std::vector<double>* out) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) {
        set_error("cannot open " + path);
        return MOAO_ERR_IO;
    }
    std::map<std::string, std::string> keys;
    char card[81] = {0};
    size_t nread = 0;
    auto fail = [&](const std::string& why) {
        std::fclose(f);
        set_error(path + ": " + why);
        return MOAO_ERR_IO;
    };
    for (;;) {
        if (std::fread(card, 1, 80, f) != 80) return fail("truncated FITS header");
        nread += 80;
        std::string c(card, 80), k = c.substr(0, 8);
        k.erase(k.find_last_not_of(' ') + 1);
        if (k == "END") break;
        if (c.compare(8, 2, "= ") == 0) {
            std::string v = c.substr(10);
            v = v.substr(0, v.find('/'));
            v.erase(0, v.find_first_not_of(' '));
            v.erase(v.find_last_not_of(' ') + 1);
            keys[k] = v;
        }
    }
    if (keys["SIMPLE"] != "T") return fail("not a FITS primary HDU");
    if ((keys.count("BZERO") && std::atof(keys["BZERO"].c_str()) != 0.0) ||
        (keys.count("BSCALE") && std::atof(keys["BSCALE"].c_str()) != 1.0))
        return fail("scaled images are not supported");
    const int bitpix = std::atoi(keys["BITPIX"].c_str());
    if (bitpix != -64 && bitpix != -32 && bitpix != 8 && bitpix != 16 && bitpix != 32 &&
        bitpix != 64)
        return fail("unsupported BITPIX " + keys["BITPIX"]);
    const int naxis = std::atoi(keys["NAXIS"].c_str());
    if (naxis < 1 || naxis > 2) return fail("only 1-D and 2-D images are supported");
    *nx = std::atol(keys["NAXIS1"].c_str());
    *ny = naxis == 2 ? std::atol(keys["NAXIS2"].c_str()) : 1;
    if (*nx < 0 || *ny < 0) return fail("negative axis length");
    if (!out) {
        std::fclose(f);
        return MOAO_OK;
    }
    std::fseek(f, (long)((nread + 2879) / 2880 * 2880), SEEK_SET);
    const size_t n = (size_t)(*nx) * (*ny), w = std::abs(bitpix) / 8;
    std::vector<unsigned char> raw(n * w);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) return fail("truncated data");
    std::fclose(f);
    out->resize(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t u = 0;
        for (size_t b = 0; b < w; ++b) u = (u << 8) | raw[i * w + b];  // big-endian
        double v;
        if (bitpix == -64) {
            std::memcpy(&v, &u, 8);
        } else if (bitpix == -32) {
            const uint32_t u32 = (uint32_t)u;
            float fv;
            std::memcpy(&fv, &u32, 4);
            v = fv;
        } else if (bitpix == 8) {
            v = (double)u;
        } else {  // signed integers
            const int sh = 64 - 8 * (int)w;
            v = (double)((int64_t)(u << sh) >> sh);
        }
        (*out)[i] = v;
    }
    return MOAO_OK;
}

int fits_write(const std::string& path, long nx, long ny, const double* data) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) {
        set_error("cannot write " + path);
        return MOAO_ERR_IO;
    }
    std::string hdr;
    auto card = [&](const std::string& k, const std::string& v) {
        char c[81];
        std::snprintf(c, sizeof c, "%-8s= %20s", k.c_str(), v.c_str());
        std::string s(c);
        s.resize(80, ' ');
        hdr += s;
    };
    card("SIMPLE", "T");
    card("BITPIX", "-64");
    card("NAXIS", "2");
    card("NAXIS1", std::to_string(nx));
    card("NAXIS2", std::to_string(ny));
    card("EXTEND", "T");
    std::string end = "END";
    end.resize(80, ' ');
    hdr += end;
    hdr.resize((hdr.size() + 2879) / 2880 * 2880, ' ');
    bool ok = std::fwrite(hdr.data(), 1, hdr.size(), f) == hdr.size();
    const size_t n = (size_t)nx * ny;
    std::vector<uint64_t> be(n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t u;
        std::memcpy(&u, data + i, 8);
        be[i] = __builtin_bswap64(u);
    }
    ok = ok && std::fwrite(be.data(), 8, n, f) == n;
    const size_t pad = (2880 - (n * 8) % 2880) % 2880;
    std::vector<char> z(pad, 0);
    ok = ok && std::fwrite(z.data(), 1, pad, f) == pad;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        set_error("write failed: " + path);
        return MOAO_ERR_IO;
    }
    return MOAO_OK;
}

}  // namespace moao

extern "C" {

int moao_fits_read(const char* path, long* nx, long* ny, double* out, long cap) {
    if (!path || !nx || !ny) {
        moao::set_error("moao_fits_read: null argument");
        return MOAO_ERR_ARG;
    }
    std::vector<double> v;
    long a = 0, b = 0;
    if (int rc = moao::fits_read(path, &a, &b, out ? &v : nullptr)) return rc;
    *nx = a;
    *ny = b;
    if (out) {
        if ((long)v.size() > cap) {
            moao::set_error(std::string("moao_fits_read: ") + path + " holds " +
                            std::to_string(v.size()) + " values, the buffer " +
                            std::to_string(cap));
            return MOAO_ERR_ARG;
        }
        std::memcpy(out, v.data(), v.size() * sizeof(double));
    }
    return MOAO_OK;
}

int moao_fits_write(const char* path, long nx, long ny, const double* data) {
    if (!path || !data || nx < 0 || ny < 0) {
        moao::set_error("moao_fits_write: invalid argument");
        return MOAO_ERR_ARG;
    }
    return moao::fits_write(path, nx, ny, data);
}

}  // extern "C"
