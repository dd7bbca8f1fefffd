// Offset and cull products on disk (see speckle_io.h).  Host code only.
//
// Raw rasters are written in chunks through a byte-swapping staging
// buffer; MSB float32 is the big-endian encoding of each value, which is
// what the reference's fwriteBS produces on a little-endian host
// (Strack/sTrackOut.c:11-15).  Text headers use the reference's printf
// formats; the VRT files are GDAL VRTRawRasterBand descriptions with the
// metadata items the reference records, in the order it inserts them.
#include <ctype.h>
#include <libgen.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <exception>
#include <string>
#include <vector>

#include "speckle_io.h"

namespace {

thread_local std::string g_err;

struct IoError {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw IoError{code, msg}; }

template <class F>
int io_guard(F &&f)
{
    try {
        f();
        return ST_OK;
    } catch (const IoError &e) {
        g_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return ST_ERR_IO;
    }
}

// printf("%s", s) of the reference, where a missing string prints as glibc does
std::string str_or_null(const char *s) { return s ? std::string(s) : std::string("(null)"); }

std::string fmt(const char *f, ...) __attribute__((format(printf, 1, 2)));
std::string fmt(const char *f, ...)
{
    char buf[4096];
    va_list ap;
    va_start(ap, f);
    vsnprintf(buf, sizeof buf, f, ap);
    va_end(ap);
    return buf;
}

std::string dir_of(const std::string &path)
{
    std::vector<char> b(path.begin(), path.end());
    b.push_back('\0');
    return dirname(b.data());
}

std::string base_of(const std::string &path)
{
    std::vector<char> b(path.begin(), path.end());
    b.push_back('\0');
    return basename(b.data());
}

struct File {
    FILE *fp = nullptr;
    std::string name;
    File(const std::string &n, const char *mode) : name(n)
    {
        fp = fopen(n.c_str(), mode);
        if (!fp)
            fail(ST_ERR_IO, "cannot open " + n);
    }
    ~File()
    {
        if (fp)
            fclose(fp);
    }
    void close()
    {
        if (fp && fclose(fp) != 0) {
            fp = nullptr;
            fail(ST_ERR_IO, "error closing " + name);
        }
        fp = nullptr;
    }
};

void write_f32(const std::string &name, const float *v, size_t n, int order)
{
    File f(name, "wb");
    constexpr size_t kChunk = 1 << 16;
    std::vector<uint32_t> buf(kChunk);
    for (size_t o = 0; o < n; o += kChunk) {
        const size_t m = std::min(kChunk, n - o);
        memcpy(buf.data(), v + o, m * 4);
        if (order == ST_MSB)
            for (size_t k = 0; k < m; k++)
                buf[k] = __builtin_bswap32(buf[k]);
        if (fwrite(buf.data(), 4, m, f.fp) != m)
            fail(ST_ERR_IO, "short write to " + name);
    }
    f.close();
}

void write_u8(const std::string &name, const uint8_t *v, size_t n)
{
    File f(name, "wb");
    if (n && fwrite(v, 1, n, f.fp) != n)
        fail(ST_ERR_IO, "short write to " + name);
    f.close();
}

void read_f32(const std::string &name, float *v, size_t n, int order)
{
    File f(name, "rb");
    if (fread(v, 4, n, f.fp) != n)
        fail(ST_ERR_IO, "short read from " + name);
    if (order == ST_MSB) {
        uint32_t *u = reinterpret_cast<uint32_t *>(v);
        for (size_t k = 0; k < n; k++)
            u[k] = __builtin_bswap32(u[k]);
    }
}

void read_u8(const std::string &name, uint8_t *v, size_t n)
{
    File f(name, "rb");
    if (n && fread(v, 1, n, f.fp) != n)
        fail(ST_ERR_IO, "short read from " + name);
}

std::string xml_escape(const std::string &s)
{
    std::string o;
    for (char c : s) {
        switch (c) {
        case '&': o += "&amp;"; break;
        case '<': o += "&lt;"; break;
        case '>': o += "&gt;"; break;
        case '"': o += "&quot;"; break;
        default: o += c;
        }
    }
    return o;
}

struct Band {
    std::string file, name;
    bool f32;
};

// A band file beside the VRT is named relative to it; one elsewhere keeps
// the path it was given.
std::string source_element(const std::string &vrtFile, const std::string &file)
{
    if (dir_of(file) == dir_of(vrtFile))
        return "    <SourceFilename relativeToVRT=\"1\">" + xml_escape(base_of(file)) + "</SourceFilename>\n";
    return "    <SourceFilename relativeToVRT=\"0\">" + xml_escape(file) + "</SourceFilename>\n";
}

typedef std::vector<std::pair<std::string, std::string>> Meta;

// One GDAL VRT of raw rasters, all nA x nR, side by side with vrtFile.
void write_vrt(const std::string &vrtFile, int32_t nR, int32_t nA, const Meta &meta, const std::vector<Band> &bands,
               int order, bool noData)
{
    std::string x = fmt("<VRTDataset rasterXSize=\"%d\" rasterYSize=\"%d\">\n", nR, nA);
    x += "  <Metadata>\n";
    for (const auto &kv : meta)
        x += "    <MDI key=\"" + xml_escape(kv.first) + "\">" + xml_escape(kv.second) + "</MDI>\n";
    x += "  </Metadata>\n";
    for (size_t b = 0; b < bands.size(); b++) {
        const Band &B = bands[b];
        x += fmt("  <VRTRasterBand dataType=\"%s\" band=\"%d\" subClass=\"VRTRawRasterBand\">\n",
                 B.f32 ? "Float32" : "Byte", (int)b + 1);
        x += "    <Description>" + xml_escape(B.name) + "</Description>\n";
        if (noData)
            x += "    <NoDataValue>-2000000000</NoDataValue>\n";
        x += source_element(vrtFile, B.file);
        x += "    <ImageOffset>0</ImageOffset>\n";
        x += fmt("    <PixelOffset>%d</PixelOffset>\n", B.f32 ? 4 : 1);
        x += fmt("    <LineOffset>%lld</LineOffset>\n", (long long)nR * (B.f32 ? 4 : 1));
        if (B.f32)
            x += std::string("    <ByteOrder>") + (order == ST_MSB ? "MSB" : "LSB") + "</ByteOrder>\n";
        x += "  </VRTRasterBand>\n";
    }
    x += "</VRTDataset>\n";
    File f(vrtFile, "w");
    if (fwrite(x.data(), 1, x.size(), f.fp) != x.size())
        fail(ST_ERR_IO, "short write to " + vrtFile);
    f.close();
}

void check_grid(int32_t nR, int32_t nA, int32_t order)
{
    if (nR <= 0 || nA <= 0)
        fail(ST_ERR_ARG, "grid size");
    if (order != ST_MSB && order != ST_LSB)
        fail(ST_ERR_ARG, "byte order");
}

}  // namespace

// ---------------------------------------------------------------------
// strack parameter file (parseTrack, Strack/parseTrack.c:13-190)

namespace {

struct RdfKeys {
    std::vector<std::pair<std::string, std::string>> kv;  // lower-case keyword, value
    const char *value(const char *key) const
    {
        std::string k(key);
        for (auto &c : k)
            c = (char)tolower((unsigned char)c);
        for (const auto &p : kv)
            if (p.first == k)
                return p.second.c_str();
        return nullptr;
    }
};

std::string trim(const std::string &x)
{
    size_t a = 0, b = x.size();
    while (a < b && isspace((unsigned char)x[a]))
        a++;
    while (b > a && isspace((unsigned char)x[b - 1]))
        b--;
    return x.substr(a, b - a);
}

// "keyword (units) = value ! comment" lines; others are skipped
RdfKeys rdf_read(const char *path)
{
    FILE *f = fopen(path, "r");
    if (!f)
        fail(ST_ERR_IO, fmt("parseTrack: cannot open %s", path));
    RdfKeys r;
    char line[8192];
    while (fgets(line, sizeof line, f)) {
        std::string l(line);
        const size_t bang = l.find('!');
        if (bang != std::string::npos)
            l.resize(bang);
        const size_t eq = l.find('=');
        if (eq == std::string::npos)
            continue;
        std::string key = l.substr(0, eq);
        const size_t par = key.find('(');
        if (par != std::string::npos)
            key.resize(par);
        key = trim(key);
        if (key.empty())
            continue;
        for (auto &c : key)
            c = (char)tolower((unsigned char)c);
        r.kv.emplace_back(key, trim(l.substr(eq + 1)));
    }
    fclose(f);
    return r;
}

void put(char *dst, const char *src, const char *what)
{
    if (!src) {
        dst[0] = '\0';
        return;
    }
    if (strlen(src) >= ST_PAR_PATH)
        fail(ST_ERR_ARG, fmt("parseTrack: %s longer than %d characters", what, ST_PAR_PATH - 1));
    strcpy(dst, src);
}

bool file_exists(const std::string &p)
{
    FILE *f = fopen(p.c_str(), "rb");
    if (!f)
        return false;
    fclose(f);
    return true;
}

// the VRT of an image: the file itself when it is one, else "<image>.vrt"
bool image_vrt(const char *image, std::string &vrt)
{
    const std::string im(image);
    if (im.size() > 4 && im.compare(im.size() - 4, 4, ".vrt") == 0 && file_exists(im)) {
        vrt = im;
        return true;
    }
    if (file_exists(im + ".vrt")) {
        vrt = im + ".vrt";
        return true;
    }
    return false;
}

// parseValues (parseTrack.c:245-254) for one %i value
void int_value(const RdfKeys &k, const char *key, int32_t *dst)
{
    const char *v = k.value(key);
    if (!v)
        fail(ST_ERR_ARG, fmt("%s not specified", key));
    int x;
    if (sscanf(v, "%i", &x) != 1)
        fail(ST_ERR_ARG, fmt("Invalid %s", key));
    *dst = x;
}

// an optional %i value: kept when absent, `fallback` when unreadable
void opt_int(const RdfKeys &k, const char *key, int32_t *dst, int32_t fallback)
{
    const char *v = k.value(key);
    int x;
    if (v && sscanf(v, "%i", &x) == 1)
        *dst = x;
    else if (v)
        *dst = fallback;
}

}  // namespace

extern "C" int st_read_track_par(const char *parFile, int32_t noComplex, st_track_par_file *out)
{
    return io_guard([&] {
        if (!parFile || !out)
            fail(ST_ERR_ARG, "st_read_track_par: null argument");
        const RdfKeys k = rdf_read(parFile);
        st_track_par_file o;
        memset(&o, 0, sizeof o);
        std::string vrt;
        // images, their par files or VRTs (parseTrack.c:26-52)
        const char *im1 = k.value("image1");
        if (!im1)
            fail(ST_ERR_ARG, "parseTrack: Missing image1");
        put(o.image1, im1, "image1");
        if (image_vrt(im1, vrt)) {
            put(o.vrt1, vrt.c_str(), "image1 vrt");
        } else {
            const char *p1 = k.value("image1par");
            if (!p1)
                fail(ST_ERR_ARG, "parseTrack: Missing image1par and no vrt");
            put(o.par1, p1, "image1par");
        }
        const char *im2 = k.value("image2");
        if (!im2)
            fail(ST_ERR_ARG, "parseTrack: Missing image2");
        put(o.image2, im2, "image2");
        if (image_vrt(im2, vrt)) {
            put(o.vrt2, vrt.c_str(), "image2 vrt");
        } else {
            const char *p2 = k.value("image2par");
            if (!p2)
                fail(ST_ERR_ARG, "parseTrack: Missing image2par and no vrt");
            put(o.par2, p2, "image2par");
        }
        // mask (parseTrack.c:55-86)
        if (const char *m = k.value("maskfile")) {
            o.maskFlag = 1;
            o.maskType = ST_GEODATMASK;
            put(o.maskFile, m, "maskfile");
            const char *g = k.value("maskgeodat");
            if (!g)
                fail(ST_ERR_ARG, "parseTrack: Missing maskgeodat");
            put(o.maskGeodat, g, "maskgeodat");
        } else {
            if (const char *om = k.value("offsetmaskfile")) {
                o.maskFlag = 1;
                o.maskType = ST_OFFSETMASK;
                put(o.maskFile, om, "offsetmaskfile");
                put(o.maskGeodat, k.value("offsetmaskdat"), "offsetmaskdat");
            }
            put(o.maskVrt, k.value("offsetmaskvrt"), "offsetmaskvrt");
        }
        // interferogram and baseline (parseTrack.c:89-109)
        if (const char *in = k.value("intfile")) {
            o.intFlag = 1;
            put(o.intFile, in, "intfile");
            const char *g = k.value("intgeodat");
            if (!g)
                fail(ST_ERR_ARG, "parseTrack: Missing intgeodat");
            put(o.intGeodat, g, "intgeodat");
        }
        if (const char *bp = k.value("baseparams")) {
            o.baseFlag = 1;
            put(o.baseParams, bp, "baseparams");
            const char *g = k.value("intgeodat");
            if (!g)
                fail(ST_ERR_ARG, "parseTrack: Missing intgeodat");
            put(o.intGeodat, g, "intgeodat");
        }
        if (noComplex)
            o.baseFlag = 0;
        // initial offsets (parseTrack.c:111-124)
        if (const char *is = k.value("initShift")) {
            o.polyShift = 0;
            put(o.initShift, is, "initShift");
        } else {
            o.polyShift = 1;
            const char *io = k.value("initialoffsetfile");
            if (!io)
                fail(ST_ERR_ARG, "parseTrack: Missing initialoffsetfile");
            put(o.initialOffsetFile, io, "initialoffsetfile");
        }
        const char *of = k.value("outputfile");
        if (!of)
            fail(ST_ERR_ARG, "parseTrack: missing ouput file");
        put(o.outputFile, of, "outputfile");
        // grid and windows (parseTrack.c:160-199)
        int_value(k, "rstart", &o.rStart);
        int_value(k, "astart", &o.aStart);
        int_value(k, "deltar", &o.deltaR);
        int_value(k, "deltaa", &o.deltaA);
        int_value(k, "nR", &o.nR);
        int_value(k, "nA", &o.nA);
        int_value(k, "wr", &o.wR);
        int_value(k, "wa", &o.wA);
        int_value(k, "wra", &o.wRa);
        int_value(k, "waa", &o.wAa);
        o.edgePad = 0;
        opt_int(k, "edgepad", &o.edgePad, 0);
        o.edgePadR = o.edgePad;
        o.edgePadA = o.edgePad;
        o.scaleFactor = 1;
        opt_int(k, "scalefactor", &o.scaleFactor, 1);
        opt_int(k, "edgeR", &o.edgePadR, o.edgePad);
        if (const char *ea = k.value("edgeA")) {
            int x;
            if (sscanf(ea, "%i", &x) == 1)
                o.edgePadA = x;
            else
                o.edgePadR = o.edgePad;  // the reference resets the range pad here
        }
        int_value(k, "navgR", &o.navgR);
        int_value(k, "navgA", &o.navgA);
        if (o.scaleFactor == 0)
            fail(ST_ERR_ARG, "parseTrack: scalefactor 0");
        o.rStart /= o.scaleFactor;
        o.aStart /= o.scaleFactor;
        o.deltaR /= o.scaleFactor;
        o.deltaA /= o.scaleFactor;
        *out = o;
    });
}

extern "C" const char *st_io_last_error(void) { return g_err.c_str(); }

extern "C" int st_write_offsets(const char *stem, const st_offsets_meta *m, const float *offR, const float *offA,
                                const float *corr, const float *azd, const uint8_t *type, int32_t nR, int32_t nA,
                                int32_t byteOrder)
{
    return io_guard([&]() {
        if (!stem || !m || !offR || !offA || !corr || !type)
            fail(ST_ERR_ARG, "null argument");
        if (!m->imageFile2 || !m->intGeodat)
            fail(ST_ERR_ARG, "imageFile2 and intGeodat are needed for the geodat names");
        check_grid(nR, nA, byteOrder);
        const std::string s(stem);
        const size_t n = (size_t)nR * nA;
        // writeOffsets (sTrackOut.c:44-60), one call per grid row in the
        // reference; the rows are contiguous, so one stream per file here
        write_f32(s + ".dr", offR, n, byteOrder);
        write_f32(s + ".da", offA, n, byteOrder);
        write_f32(s + ".cc", corr, n, byteOrder);
        if (azd)
            write_f32(s + ".azd", azd, n, byteOrder);
        write_u8(s + ".mt", type, n);
        const int sf = m->scaleFactor;
        const std::string geodat(m->intGeodat), image2(m->imageFile2);
        // writeDatFile (speckleTrack.c:1453-1468): geo2 = dirname(image 2) / geodat
        {
            const std::string dat = fmt("%i %i %i %i %i %i\n", m->rStart * sf, m->aStart * sf, nR, nA,
                                        m->deltaR * sf, m->deltaA * sf) +
                                    geodat + " " + dir_of(image2) + "/" + geodat + "\n";
            File f(s + ".dat", "w");
            if (fwrite(dat.data(), 1, dat.size(), f.fp) != dat.size())
                fail(ST_ERR_IO, "short write to " + s + ".dat");
            f.close();
        }
        // writeVrtFile (writeVrt.c:16-70): geo2 = dirname(image 2) / basename(geodat)
        Meta meta = {
            {"r0", fmt("%i", m->rStart * sf)},
            {"a0", fmt("%i", m->aStart * sf)},
            {"deltaA", fmt("%i", m->deltaA * sf)},
            {"deltaR", fmt("%i", m->deltaR * sf)},
            {"sigmaStreaks", fmt("%f", 0.)},
            {"sigmaRange", fmt("%f", 0.)},
            {"geo1", geodat},
            {"geo2", dir_of(image2) + "/" + base_of(geodat)},
            {"Image1", str_or_null(m->imageFile1)},
            {"Image2", image2},
            {"mask", str_or_null(m->maskFile)},
            {"wR", fmt("%i", m->wR)},
            {"wA", fmt("%i", m->wA)},
            {"wRa", fmt("%i", m->wRa)},
            {"wAa", fmt("%i", m->wAa)},
            {"scaleFactor", fmt("%i", sf)},
        };
        write_vrt(s + ".mt.vrt", nR, nA, meta, {{s + ".mt", "MatchType", false}}, byteOrder, false);
        meta.push_back({"ByteOrder", byteOrder == ST_LSB ? "LSB" : "MSB"});
        write_vrt(s + ".vrt", nR, nA, meta,
                  {{s + ".dr", "RangeOffsets", true}, {s + ".da", "AzimuthOffsets", true},
                   {s + ".cc", "Correlation", true}},
                  byteOrder, true);
    });
}

extern "C" int st_read_dat(const char *datFile, int32_t hdr[6], char *geo1, char *geo2, size_t geoLen)
{
    return io_guard([&]() {
        if (!datFile || !hdr)
            fail(ST_ERR_ARG, "null argument");
        File f(datFile, "r");
        char line[1024];
        // getDataString skips comment lines starting with ';'
        auto next = [&](char *buf) -> bool {
            while (fgets(buf, sizeof line, f.fp))
                if (buf[0] != ';' && buf[0] != '\n')
                    return true;
            return false;
        };
        if (!next(line) || sscanf(line, "%i%i%i%i%i%i", &hdr[0], &hdr[1], &hdr[2], &hdr[3], &hdr[4], &hdr[5]) != 6)
            fail(ST_ERR_IO, std::string("missing image parameters in ") + datFile);
        if (geo1 && geoLen)
            geo1[0] = '\0';
        if (geo2 && geoLen)
            geo2[0] = '\0';
        if (next(line) && strstr(line, "geodat")) {
            char a[1024] = {0}, b[1024] = {0};
            const int k = sscanf(line, "%1023s %1023s", a, b);
            if (geo1 && geoLen && k >= 1)
                snprintf(geo1, geoLen, "%s", a);
            if (geo2 && geoLen && k >= 2)
                snprintf(geo2, geoLen, "%s", b);
        }
    });
}

extern "C" int st_read_offsets(const char *stem, int32_t nR, int32_t nA, int32_t byteOrder, float *offR, float *offA,
                               float *corr, uint8_t *type)
{
    return io_guard([&]() {
        if (!stem || !offR || !offA || !corr || !type)
            fail(ST_ERR_ARG, "null argument");
        check_grid(nR, nA, byteOrder);
        const std::string s(stem);
        const size_t n = (size_t)nR * nA;
        read_f32(s + ".dr", offR, n, byteOrder);
        read_f32(s + ".da", offA, n, byteOrder);
        read_f32(s + ".cc", corr, n, byteOrder);
        read_u8(s + ".mt", type, n);
        // offA < (-LARGEINT + 1): LARGEINT comes from the un-vendored clib
        // headers; read as the integer 2000000000, the bound converts to
        // the float -2e9, so the -2e9 sentinel itself is not below it
        const float cut = (float)(-2000000000 + 1);
        for (size_t k = 0; k < n; k++)
            if (offA[k] < cut)
                type[k] = 0;
    });
}

extern "C" int st_write_cull(const char *outBase, const char *inBase, const st_cull_meta *m, const float *offRS,
                             const float *offAS, const float *sigmaR, const float *sigmaA, const float *corr,
                             const uint8_t *type, int32_t nR, int32_t nA, int32_t byteOrder)
{
    return io_guard([&]() {
        if (!outBase || !inBase || !m || !offRS || !offAS || !sigmaR || !sigmaA || !corr || !type)
            fail(ST_ERR_ARG, "null argument");
        check_grid(nR, nA, byteOrder);
        const std::string o(outBase), in(inBase);
        const size_t n = (size_t)nR * nA;
        // file names of cullst.c:240-285: the sigma files take the input base
        write_f32(o + ".dr", offRS, n, byteOrder);
        write_f32(o + ".da", offAS, n, byteOrder);
        write_f32(in + ".sr", sigmaR, n, byteOrder);
        write_f32(in + ".sa", sigmaA, n, byteOrder);
        write_f32(o + ".cc", corr, n, byteOrder);
        write_u8(o + ".mt", type, n);
        // populateMetaData (loadCullData.c:95-118) + insertCullParams (:38-49)
        Meta meta = {
            {"r0", fmt("%i", m->rStart)},
            {"a0", fmt("%i", m->aStart)},
            {"deltaA", fmt("%i", m->deltaA)},
            {"deltaR", fmt("%i", m->deltaR)},
            {"sigmaStreaks", fmt("%f", 0.)},
            {"sigmaRange", fmt("%f", 0.)},
            {"sR", fmt("%i", m->sR)},
            {"sA", fmt("%i", m->sA)},
            {"boxSize", fmt("%i", m->boxSize)},
            {"nGood", fmt("%i", m->nGood)},
            {"maxR", fmt("%f", (double)m->maxR)},
            {"maxA", fmt("%f", (double)m->maxA)},
            {"corrThresh", fmt("%f", (double)m->corrThresh)},
            {"singleMT", fmt("%i", m->singleMT)},
        };
        if (m->geo2) {
            meta.push_back({"geo1", str_or_null(m->geo1)});
            meta.push_back({"geo2", m->geo2});
        }
        write_vrt(o + ".mt.vrt", nR, nA, meta, {{o + ".mt", "MatchType", false}}, byteOrder, false);
        meta.push_back({"ByteOrder", byteOrder == ST_LSB ? "LSB" : "MSB"});
        // writeCullVrt (writeCullData.c:15-47)
        write_vrt(o + ".vrt", nR, nA, meta,
                  {{o + ".dr", "RangeOffsets", true}, {o + ".da", "AzimuthOffsets", true},
                   {in + ".sr", "RangeSigma", true}, {in + ".sa", "AzimuthSigma", true},
                   {o + ".cc", "Correlation", true}},
                  byteOrder, true);
    });
}

// ---- inputs before tracking ----------------------------------------------

namespace {

// getDataString: the next line that is neither empty nor a ';' comment
bool next_data_line(FILE *fp, char *buf, int len)
{
    while (fgets(buf, len, fp))
        if (buf[0] != ';' && buf[0] != '\n' && buf[0] != '\r')
            return true;
    return false;
}

void read_header6(const std::string &name, const char *fmt6, void *a, void *b, void *c, void *d, void *e, void *f)
{
    File fp(name, "r");
    char line[1024];
    if (!next_data_line(fp.fp, line, sizeof line) || sscanf(line, fmt6, a, b, c, d, e, f) != 6)
        fail(ST_ERR_IO, "missing image parameters in " + name);
}

}  // namespace

extern "C" int st_read_offset_polys(const char *file, double rShiftPoly[3], double aShiftPoly[3])
{
    return io_guard([&]() {
        if (!file || !rShiftPoly || !aShiftPoly)
            fail(ST_ERR_ARG, "null argument");
        File fp(file, "r");
        char line[1024];
        bool rDone = false, aDone = false;
        while (!rDone || !aDone) {
            if (!next_data_line(fp.fp, line, sizeof line))
                fail(ST_ERR_IO, std::string("missing polynomial in ") + file);
            if (const char *t = strstr(line, "range_offset_polynomial:")) {
                t = strchr(t, ':') + 1;
                if (sscanf(t, "%lf%lf%lf", &rShiftPoly[0], &rShiftPoly[1], &rShiftPoly[2]) != 3)
                    fail(ST_ERR_IO, std::string("missing range parameter in ") + file);
                rDone = true;
            } else if (const char *t2 = strstr(line, "azimuth_offset_polynomial:")) {
                t2 = strchr(t2, ':') + 1;
                if (sscanf(t2, "%lf%lf%lf", &aShiftPoly[0], &aShiftPoly[1], &aShiftPoly[2]) != 3)
                    fail(ST_ERR_IO, std::string("missing azimuth parameter in ") + file);
                aDone = true;
            }
        }
    });
}

extern "C" int st_read_initial_offsets_header(const char *daFile, int32_t *r0, int32_t *a0, int32_t *nr, int32_t *na,
                                              float *deltaR, float *deltaA)
{
    return io_guard([&]() {
        if (!daFile || !r0 || !a0 || !nr || !na || !deltaR || !deltaA)
            fail(ST_ERR_ARG, "null argument");
        read_header6(std::string(daFile) + ".dat", "%i%i%i%i%f%f", r0, a0, nr, na, deltaR, deltaA);
    });
}

extern "C" int st_read_initial_offsets(const char *daFile, const char *drFile, int32_t nr, int32_t na, float *da,
                                       float *dr)
{
    return io_guard([&]() {
        if (!daFile || !drFile || !da || !dr)
            fail(ST_ERR_ARG, "null argument");
        if (nr <= 0 || na <= 0)
            fail(ST_ERR_ARG, "grid size");
        read_f32(daFile, da, (size_t)nr * na, ST_MSB);
        read_f32(drFile, dr, (size_t)nr * na, ST_MSB);
    });
}

extern "C" int st_read_mask_header(const char *datFile, int32_t hdr[6])
{
    return io_guard([&]() {
        if (!datFile || !hdr)
            fail(ST_ERR_ARG, "null argument");
        float dR = 0, dA = 0;
        read_header6(datFile, "%i%i%i%i%f%f", &hdr[0], &hdr[1], &hdr[2], &hdr[3], &dR, &dA);
        hdr[4] = (int)dR;
        hdr[5] = (int)dA;
    });
}

extern "C" int st_read_mask(const char *maskFile, int32_t nr, int32_t na, uint8_t *mask)
{
    return io_guard([&]() {
        if (!maskFile || !mask)
            fail(ST_ERR_ARG, "null argument");
        if (nr <= 0 || na <= 0)
            fail(ST_ERR_ARG, "grid size");
        read_u8(maskFile, mask, (size_t)nr * na);
    });
}

extern "C" int st_read_interferogram(const char *intFile, int32_t nr, int32_t na, float *intf)
{
    return io_guard([&]() {
        if (!intFile || !intf)
            fail(ST_ERR_ARG, "null argument");
        if (nr <= 0 || na <= 0)
            fail(ST_ERR_ARG, "interferogram size");
        const size_t n = (size_t)nr * na;
        for (size_t k = 0; k < 2 * n; k++)
            intf[k] = 1.0f;
        File f(intFile, "rb");
        if (fseek(f.fp, 0, SEEK_END) != 0)
            fail(ST_ERR_IO, std::string("cannot size ") + intFile);
        const long bytes = ftell(f.fp);
        rewind(f.fp);
        const long lines = std::min<long>(bytes / ((long)nr * 4), na);
        std::vector<uint16_t> row(2 * (size_t)nr);
        for (long i = 0; i < lines; i++) {
            if (fread(row.data(), 4, (size_t)nr, f.fp) != (size_t)nr)
                fail(ST_ERR_IO, std::string("short read from ") + intFile);
            float *dst = intf + 2 * (size_t)i * nr;
            for (size_t k = 0; k < 2 * (size_t)nr; k++)
                dst[k] = (float)(int16_t)__builtin_bswap16(row[k]);
        }
    });
}

// ---- Cullls products ------------------------------------------------------

extern "C" int st_write_cull_ls(const char *matchFile, const st_cull_ls_meta *m, const float *XS, const float *YS,
                                const float *sigmaX, const float *sigmaY, const uint8_t *type, int32_t nx, int32_t ny,
                                int32_t nMatch, int32_t nAttempt, int32_t byteOrder)
{
    return io_guard([&]() {
        if (!matchFile || !m || !XS || !YS || !sigmaX || !sigmaY || !type)
            fail(ST_ERR_ARG, "null argument");
        check_grid(nx, ny, byteOrder);
        const std::string b(matchFile);
        const size_t n = (size_t)nx * ny;
        // final count and RMS sigmas (cullls.c:84-101)
        int nGoodPts = 0;
        double sx = 0, sy = 0;
        const double lim = (double)m->noData + 1.0;
        for (size_t k = 0; k < n; k++)
            if ((double)XS[k] > lim && (double)YS[k] > lim) {
                sx += sigmaX[k] * sigmaX[k];
                sy += sigmaY[k] * sigmaY[k];
                nGoodPts++;
            }
        sx = sqrt(sx / (double)std::max(nGoodPts, 1));
        sy = sqrt(sy / (double)std::max(nGoodPts, 1));
        write_f32(b + ".cull.dx", XS, n, byteOrder);
        write_f32(b + ".cull.dy", YS, n, byteOrder);
        write_f32(b + ".cull.sx", sigmaX, n, byteOrder);
        write_f32(b + ".cull.sy", sigmaY, n, byteOrder);
        write_u8(b + ".cull.mtype", type, n);
        // the header of cullls.c:150-174 (its stepY line prints stepX and
        // its stepX line stepY)
        File f(b + ".cull.dat", "w");
        std::string t;
        t += "fileEarly = " + str_or_null(m->fileEarly) + "\n";
        t += "fileLate = " + str_or_null(m->fileLate) + "\n";
        t += fmt("x0 = %11.2lf\n", m->x0);
        t += fmt("y0 = %11.2lf\n", m->y0);
        t += fmt("dx = %10.5lf\n", m->dx);
        t += fmt("dy = %10.5lf\n", m->dy);
        t += fmt("stepY = %u\n", m->stepX);
        t += fmt("stepX = %u\n", m->stepY);
        t += fmt("nx = %9i\n", nx);
        t += fmt("ny = %9i\n", ny);
        t += fmt("slowFlag = %9i\n", m->slowFlag);
        t += fmt("EPSG = %9i\n", m->proj);
        t += fmt("earlyImageJD = %10lf\n", m->jdEarly);
        t += fmt("lateImageJD = %10lf\n", m->jdLate);
        t += fmt("IntervalBetweenImages = %5i\n", (int)(m->jdLate - m->jdEarly));
        t += fmt("Success_rate_for_attempted_matches(%%) =  %7.2f \n",
                 ((float)nMatch / (float)std::max(nAttempt, 1)) * 100.);
        t += fmt("Culled_rate_for_attempted_matches(%%) =  %7.2f \n",
                 ((float)nGoodPts / (float)std::max(nAttempt, 1)) * 100.);
        t += fmt("Mean_sigmaX = %11.3lf\n", sx);
        t += fmt("Mean_sigmaY = %11.3lf\n", sy);
        t += "& \n";
        if (fwrite(t.data(), 1, t.size(), f.fp) != t.size())
            fail(ST_ERR_IO, "short write to " + b + ".cull.dat");
        f.close();
    });
}
