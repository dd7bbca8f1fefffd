// GDAL VRT files whose bands are VRTRawRasterBand (see speckle_io.h).  Host
// code only.  A small XML reader takes the elements the VRT driver defines
// for raw bands (GDAL's VRT format description): the dataset size, the
// default-domain metadata items, and per band its data type, description,
// band metadata, source file (relative to the VRT or not), image / pixel /
// line offsets and byte order.  Omitted offsets and byte
// order take GDAL's defaults: 0, the data type size, pixel offset x width,
// and the machine order (LSB here).
#include <ctype.h>
#include <libgen.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <exception>
#include <map>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "speckle_io.h"

namespace {

thread_local std::string v_err;

struct VrtError {
    int code;
    std::string msg;
};

[[noreturn]] void vfail(int code, const std::string &m) { throw VrtError{code, m}; }

template <class F>
int vguard(F &&f)
{
    try {
        f();
        return ST_OK;
    } catch (const VrtError &e) {
        v_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        v_err = e.what();
        return ST_ERR_IO;
    }
}

// ---- XML -------------------------------------------------------------------

struct Node {
    std::string name, text;
    std::map<std::string, std::string> attr;
    std::vector<Node> kids;
    const Node *child(const char *n) const
    {
        for (const Node &k : kids)
            if (k.name == n)
                return &k;
        return nullptr;
    }
};

std::string unescape(const std::string &s)
{
    std::string o;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] != '&') {
            o += s[i];
            continue;
        }
        const size_t e = s.find(';', i);
        if (e == std::string::npos)
            vfail(ST_ERR_IO, "bad XML entity");
        const std::string ent = s.substr(i + 1, e - i - 1);
        if (ent == "lt")
            o += '<';
        else if (ent == "gt")
            o += '>';
        else if (ent == "amp")
            o += '&';
        else if (ent == "quot")
            o += '"';
        else if (ent == "apos")
            o += '\'';
        else if (!ent.empty() && ent[0] == '#')
            o += (char)strtol(ent.c_str() + (ent.size() > 1 && ent[1] == 'x' ? 2 : 1), nullptr,
                              ent.size() > 1 && ent[1] == 'x' ? 16 : 10);
        else
            vfail(ST_ERR_IO, "unknown XML entity &" + ent + ";");
        i = e;
    }
    return o;
}

std::string trim(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a]))
        a++;
    while (b > a && isspace((unsigned char)s[b - 1]))
        b--;
    return s.substr(a, b - a);
}

struct Parser {
    const std::string &s;
    size_t p = 0;
    explicit Parser(const std::string &x) : s(x) {}

    void skip_misc()
    {
        while (true) {
            while (p < s.size() && isspace((unsigned char)s[p]))
                p++;
            if (s.compare(p, 4, "<!--") == 0) {
                const size_t e = s.find("-->", p);
                if (e == std::string::npos)
                    vfail(ST_ERR_IO, "unterminated XML comment");
                p = e + 3;
            } else if (s.compare(p, 2, "<?") == 0) {
                const size_t e = s.find("?>", p);
                if (e == std::string::npos)
                    vfail(ST_ERR_IO, "unterminated XML declaration");
                p = e + 2;
            } else {
                return;
            }
        }
    }
    std::string name()
    {
        const size_t a = p;
        while (p < s.size() && (isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == ':' || s[p] == '-' || s[p] == '.'))
            p++;
        if (p == a)
            vfail(ST_ERR_IO, "XML name expected");
        return s.substr(a, p - a);
    }
    Node element()
    {
        skip_misc();
        if (p >= s.size() || s[p] != '<')
            vfail(ST_ERR_IO, "XML element expected");
        p++;
        Node n;
        n.name = name();
        while (true) {
            while (p < s.size() && isspace((unsigned char)s[p]))
                p++;
            if (p >= s.size())
                vfail(ST_ERR_IO, "unterminated tag <" + n.name + ">");
            if (s[p] == '/') {
                if (s.compare(p, 2, "/>") != 0)
                    vfail(ST_ERR_IO, "bad tag end");
                p += 2;
                return n;
            }
            if (s[p] == '>') {
                p++;
                break;
            }
            const std::string a = name();
            while (p < s.size() && isspace((unsigned char)s[p]))
                p++;
            if (p >= s.size() || s[p] != '=')
                vfail(ST_ERR_IO, "attribute value expected");
            p++;
            while (p < s.size() && isspace((unsigned char)s[p]))
                p++;
            const char q = p < s.size() ? s[p] : 0;
            if (q != '"' && q != '\'')
                vfail(ST_ERR_IO, "quoted attribute value expected");
            const size_t e = s.find(q, p + 1);
            if (e == std::string::npos)
                vfail(ST_ERR_IO, "unterminated attribute value");
            n.attr[a] = unescape(s.substr(p + 1, e - p - 1));
            p = e + 1;
        }
        // content: text and child elements up to the end tag
        while (true) {
            const size_t lt = s.find('<', p);
            if (lt == std::string::npos)
                vfail(ST_ERR_IO, "missing </" + n.name + ">");
            n.text += unescape(s.substr(p, lt - p));
            p = lt;
            if (s.compare(p, 4, "<!--") == 0) {
                skip_misc();
                continue;
            }
            if (s.compare(p, 2, "</") == 0) {
                p += 2;
                const std::string en = name();
                if (en != n.name)
                    vfail(ST_ERR_IO, "</" + en + "> closes <" + n.name + ">");
                while (p < s.size() && isspace((unsigned char)s[p]))
                    p++;
                if (p >= s.size() || s[p] != '>')
                    vfail(ST_ERR_IO, "bad end tag");
                p++;
                n.text = trim(n.text);
                return n;
            }
            n.kids.push_back(element());
        }
    }
};

// ---- raw bands -------------------------------------------------------------

struct TypeInfo {
    const char *name;
    int size;
    bool cplx;
};

const TypeInfo kTypes[] = {{"Byte", 1, false},    {"Int16", 2, false},   {"UInt16", 2, false},
                           {"Int32", 4, false},   {"UInt32", 4, false},  {"Float32", 4, false},
                           {"Float64", 8, false}, {"CInt16", 4, true},   {"CFloat32", 8, true}};

const TypeInfo *type_info(const std::string &n)
{
    for (const TypeInfo &t : kTypes)
        if (n == t.name)
            return &t;
    return nullptr;
}

long long to_ll(const std::string &s, const char *what)
{
    char *e = nullptr;
    const long long v = strtoll(s.c_str(), &e, 10);
    if (e == s.c_str() || *trim(e).c_str())
        vfail(ST_ERR_IO, std::string("bad ") + what + ": " + s);
    return v;
}

}  // namespace

struct st_vrt {
    std::string file;
    int32_t nx = 0, ny = 0;
    std::vector<std::pair<std::string, std::string>> meta;
    struct Band {
        const TypeInfo *type;
        std::string desc, src;
        std::vector<std::pair<std::string, std::string>> meta;
        long long imageOffset, pixelOffset, lineOffset;
        bool msb;
    };
    std::vector<Band> bands;

    const Band &band(int32_t b) const
    {
        if (b < 1 || b > (int32_t)bands.size())
            vfail(ST_ERR_ARG, "band " + std::to_string(b) + " out of range");
        return bands[(size_t)b - 1];
    }

    // One row of band b, each pixel's bytes in the machine order.
    template <class Emit>
    void read_rows(const Band &B, Emit emit) const
    {
        FILE *fp = fopen(B.src.c_str(), "rb");
        if (!fp)
            vfail(ST_ERR_IO, "cannot open " + B.src);
        const int es = B.type->cplx ? B.type->size / 2 : B.type->size;
        const size_t span = (size_t)(nx - 1) * (size_t)B.pixelOffset + (size_t)B.type->size;
        std::vector<unsigned char> row(span), px((size_t)B.type->size);
        try {
            for (int32_t y = 0; y < ny; y++) {
                if (fseeko(fp, (off_t)(B.imageOffset + (long long)y * B.lineOffset), SEEK_SET) != 0 ||
                    fread(row.data(), 1, span, fp) != span)
                    vfail(ST_ERR_IO, "short read from " + B.src);
                for (int32_t x = 0; x < nx; x++) {
                    memcpy(px.data(), row.data() + (size_t)x * B.pixelOffset, (size_t)B.type->size);
                    if (B.msb && es > 1)
                        for (int c = 0; c < B.type->size; c += es)
                            for (int k = 0; k < es / 2; k++)
                                std::swap(px[(size_t)c + k], px[(size_t)c + es - 1 - k]);
                    emit((size_t)y * nx + x, px.data());
                }
            }
        } catch (...) {
            fclose(fp);
            throw;
        }
        fclose(fp);
    }
};

namespace {

double real_value(const TypeInfo *t, const unsigned char *p)
{
    const std::string n = t->name;
    if (n == "Byte")
        return p[0];
    if (n == "Int16") {
        int16_t v;
        memcpy(&v, p, 2);
        return v;
    }
    if (n == "UInt16") {
        uint16_t v;
        memcpy(&v, p, 2);
        return v;
    }
    if (n == "Int32") {
        int32_t v;
        memcpy(&v, p, 4);
        return v;
    }
    if (n == "UInt32") {
        uint32_t v;
        memcpy(&v, p, 4);
        return v;
    }
    if (n == "Float32") {
        float v;
        memcpy(&v, p, 4);
        return v;
    }
    double v;
    memcpy(&v, p, 8);
    return v;
}

void copy_str(const std::string &s, char *out, size_t len)
{
    if (!out || len == 0)
        vfail(ST_ERR_ARG, "null output buffer");
    if (s.size() + 1 > len)
        vfail(ST_ERR_ARG, "output buffer too small");
    memcpy(out, s.c_str(), s.size() + 1);
}

std::vector<std::pair<std::string, std::string>> metadata_items(const Node *md)
{
    std::vector<std::pair<std::string, std::string>> out;
    if (!md)
        return out;
    for (const Node &k : md->kids)
        if (k.name == "MDI" && k.attr.count("key"))
            out.push_back({k.attr.at("key"), k.text});
    return out;
}

}  // namespace

extern "C" const char *st_vrt_last_error(void) { return v_err.c_str(); }

extern "C" int st_vrt_open(const char *vrtFile, st_vrt **out)
{
    return vguard([&]() {
        if (!vrtFile || !out)
            vfail(ST_ERR_ARG, "null argument");
        *out = nullptr;
        FILE *fp = fopen(vrtFile, "rb");
        if (!fp)
            vfail(ST_ERR_IO, std::string("cannot open ") + vrtFile);
        std::string text;
        char buf[65536];
        size_t n;
        while ((n = fread(buf, 1, sizeof buf, fp)) > 0)
            text.append(buf, n);
        fclose(fp);
        Parser P(text);
        const Node root = P.element();
        if (root.name != "VRTDataset")
            vfail(ST_ERR_IO, std::string(vrtFile) + " is not a VRTDataset");
        auto v = std::make_unique<st_vrt>();
        v->file = vrtFile;
        if (!root.attr.count("rasterXSize") || !root.attr.count("rasterYSize"))
            vfail(ST_ERR_IO, "VRTDataset without rasterXSize / rasterYSize");
        v->nx = (int32_t)to_ll(root.attr.at("rasterXSize"), "rasterXSize");
        v->ny = (int32_t)to_ll(root.attr.at("rasterYSize"), "rasterYSize");
        if (v->nx <= 0 || v->ny <= 0)
            vfail(ST_ERR_IO, "raster size");
        for (const Node &k : root.kids)
            if (k.name == "Metadata" && (!k.attr.count("domain") || k.attr.at("domain").empty()))
                for (auto &kv : metadata_items(&k))
                    v->meta.push_back(kv);
        std::vector<char> dirBuf(v->file.begin(), v->file.end());
        dirBuf.push_back('\0');
        const std::string dir = dirname(dirBuf.data());
        std::map<int, st_vrt::Band> byIndex;
        int next = 1;
        for (const Node &k : root.kids) {
            if (k.name != "VRTRasterBand")
                continue;
            if (!k.attr.count("subClass") || k.attr.at("subClass") != "VRTRawRasterBand")
                vfail(ST_ERR_UNSUPPORTED, "only VRTRawRasterBand bands are read");
            st_vrt::Band B;
            B.type = type_info(k.attr.count("dataType") ? k.attr.at("dataType") : "Byte");
            if (!B.type)
                vfail(ST_ERR_UNSUPPORTED, "data type " + k.attr.at("dataType"));
            const int idx = k.attr.count("band") ? (int)to_ll(k.attr.at("band"), "band") : next;
            next = idx + 1;
            if (const Node *d = k.child("Description"))
                B.desc = d->text;
            for (const Node &kk : k.kids)
                if (kk.name == "Metadata" && (!kk.attr.count("domain") || kk.attr.at("domain").empty()))
                    for (auto &kv : metadata_items(&kk))
                        B.meta.push_back(kv);
            const Node *src = k.child("SourceFilename");
            if (!src || src->text.empty())
                vfail(ST_ERR_IO, "raw band without SourceFilename");
            const bool rel = src->attr.count("relativeToVRT") && src->attr.at("relativeToVRT") == "1";
            B.src = rel ? dir + "/" + src->text : src->text;
            const Node *io = k.child("ImageOffset"), *po = k.child("PixelOffset"), *lo = k.child("LineOffset");
            B.imageOffset = io ? to_ll(io->text, "ImageOffset") : 0;
            B.pixelOffset = po ? to_ll(po->text, "PixelOffset") : B.type->size;
            B.lineOffset = lo ? to_ll(lo->text, "LineOffset") : B.pixelOffset * v->nx;
            if (B.imageOffset < 0 || B.pixelOffset < B.type->size || B.lineOffset < 0)
                vfail(ST_ERR_UNSUPPORTED, "band offsets");
            const Node *bo = k.child("ByteOrder");
            const std::string order = bo ? bo->text : "LSB";
            if (order != "MSB" && order != "LSB")
                vfail(ST_ERR_IO, "ByteOrder " + order);
            B.msb = order == "MSB";
            byIndex[idx] = B;
        }
        for (int i = 1; i <= (int)byIndex.size(); i++) {
            if (!byIndex.count(i))
                vfail(ST_ERR_IO, "band numbers are not 1.." + std::to_string(byIndex.size()));
            v->bands.push_back(byIndex[i]);
        }
        *out = v.release();
    });
}

extern "C" void st_vrt_close(st_vrt *v) { delete v; }

extern "C" int st_vrt_size(const st_vrt *v, int32_t *nx, int32_t *ny, int32_t *nBands)
{
    return vguard([&]() {
        if (!v || !nx || !ny || !nBands)
            vfail(ST_ERR_ARG, "null argument");
        *nx = v->nx;
        *ny = v->ny;
        *nBands = (int32_t)v->bands.size();
    });
}

extern "C" int st_vrt_metadata(const st_vrt *v, const char *key, char *value, size_t len)
{
    return vguard([&]() {
        if (!v || !key)
            vfail(ST_ERR_ARG, "null argument");
        for (const auto &kv : v->meta)
            if (kv.first == key) {
                copy_str(kv.second, value, len);
                return;
            }
        vfail(ST_ERR_ARG, std::string("no metadata item ") + key);
    });
}

extern "C" int st_vrt_band_description(const st_vrt *v, int32_t band, char *desc, size_t len)
{
    return vguard([&]() {
        if (!v)
            vfail(ST_ERR_ARG, "null argument");
        const st_vrt::Band &B = v->band(band);
        std::string d = B.desc;
        if (d.empty())
            for (const auto &kv : B.meta)
                if (kv.first == "Description")
                    d = kv.second;
        copy_str(d, desc, len);
    });
}

extern "C" int st_vrt_read_f32(const st_vrt *v, int32_t band, float *out)
{
    return vguard([&]() {
        if (!v || !out)
            vfail(ST_ERR_ARG, "null argument");
        const st_vrt::Band &B = v->band(band);
        if (B.type->cplx)
            vfail(ST_ERR_ARG, "complex band: use st_vrt_read_c64");
        v->read_rows(B, [&](size_t i, const unsigned char *p) { out[i] = (float)real_value(B.type, p); });
    });
}

extern "C" int st_vrt_read_u8(const st_vrt *v, int32_t band, uint8_t *out)
{
    return vguard([&]() {
        if (!v || !out)
            vfail(ST_ERR_ARG, "null argument");
        const st_vrt::Band &B = v->band(band);
        if (std::string(B.type->name) != "Byte")
            vfail(ST_ERR_ARG, "not a Byte band");
        v->read_rows(B, [&](size_t i, const unsigned char *p) { out[i] = p[0]; });
    });
}

extern "C" int st_vrt_read_c64(const st_vrt *v, int32_t band, float *out)
{
    return vguard([&]() {
        if (!v || !out)
            vfail(ST_ERR_ARG, "null argument");
        const st_vrt::Band &B = v->band(band);
        if (!B.type->cplx)
            vfail(ST_ERR_ARG, "not a complex band");
        const bool i16 = std::string(B.type->name) == "CInt16";
        v->read_rows(B, [&](size_t i, const unsigned char *p) {
            if (i16) {
                int16_t re, im;
                memcpy(&re, p, 2);
                memcpy(&im, p + 2, 2);
                out[2 * i] = re;
                out[2 * i + 1] = im;
            } else {
                memcpy(out + 2 * i, p, 8);
            }
        });
    });
}
