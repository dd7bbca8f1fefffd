/*
 * speckle_io.h -- the files either side of the tracking path: offset
 * products written after strack / strackw, the cullst inputs read from
 * them, and the cullst products.  Host code only (no GPU call); part of
 * libspeckle_amd.so.
 *
 * Raw rasters are row-major nA x nR: float32 (.dr .da .cc .azd .sr .sa) in
 * the requested byte order (ST_MSB, the reference default, or ST_LSB) and
 * bytes (.mt).  Text headers (.dat) use the reference's printf formats.
 * VRT files are GDAL VRTRawRasterBand descriptions of the same rasters.
 *
 * Every function returns ST_OK or a negative ST_ERR_* code;
 * st_io_last_error() gives the message of the calling thread's last
 * failure.
 */
#ifndef SPECKLE_IO_H
#define SPECKLE_IO_H

#include <stddef.h>
#include <stdint.h>

#include "speckle_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ST_ERR_IO (-6) /* file could not be opened, read or written */

/* byte orders: ST_LSB / ST_MSB of speckle_amd.h (Cullst/writeCullData.c:8-9) */

/* What the offset headers record (TrackParams fields, Strack/strack.h). */
typedef struct st_offsets_meta {
    int32_t rStart, aStart, deltaR, deltaA, scaleFactor;
    int32_t wR, wA, wRa, wAa;
    const char *intGeodat;  /* geodat of image 1 */
    const char *imageFile1, *imageFile2, *maskFile;
} st_offsets_meta;

const char *st_io_last_error(void);

/* stem.dr .da .cc (.azd when azd != NULL) .mt, stem.dat, stem.vrt (bands
 * RangeOffsets, AzimuthOffsets, Correlation, no-data -2e9, byteOrder) and
 * stem.mt.vrt (MatchType).  Replaces sTrackOut / writeOffsets
 * (Strack/sTrackOut.c:11-60), writeDatFile (Strack/speckleTrack.c:1453-1468,
 * Strackw/corrTrackFast.c:35) and writeVrtFile (Strack/writeVrt.c:16-70).
 * strackw passes azd = NULL (it writes no .azd). */
int st_write_offsets(const char *stem, const st_offsets_meta *meta, const float *offR, const float *offA,
                     const float *corr, const float *azd, const uint8_t *type, int32_t nR, int32_t nA,
                     int32_t byteOrder);

/* First line of a .dat header ("%i%i%i%i%i%i": r0 a0 nr na deltaR deltaA)
 * into hdr[6]; when the second line names geodat files, its two words into
 * geo1 / geo2 (each geoLen bytes, "" otherwise).  loadCullData's .dat path
 * (Cullst/loadCullData.c:154-168). */
int st_read_dat(const char *datFile, int32_t hdr[6], char *geo1, char *geo2, size_t geoLen);

/* stem.dr .da .cc .mt (nA x nR, byteOrder) into caller arrays, with the
 * loader's rule that a point whose azimuth offset is below -LARGEINT + 1
 * has match type 0 (Cullst/loadCullData.c:184-197). */
int st_read_offsets(const char *stem, int32_t nR, int32_t nA, int32_t byteOrder, float *offR, float *offA,
                    float *corr, uint8_t *type);

/* What the cullst product headers record (CullParams, Cullst/cullst.h). */
typedef struct st_cull_meta {
    int32_t rStart, aStart, deltaR, deltaA;
    int32_t sR, sA, boxSize, nGood, singleMT;
    float maxR, maxA, corrThresh;
    const char *geo1, *geo2; /* NULL: not recorded */
} st_cull_meta;

/* writeCullData (Cullst/writeCullData.c:15-81) with the file names of
 * cullst.c:240-300: outBase.dr .da .cc .mt, inBase.sr .sa (the reference
 * names the sigma files after the input), outBase.vrt (RangeOffsets,
 * AzimuthOffsets, RangeSigma, AzimuthSigma, Correlation) and
 * outBase.mt.vrt.  Header metadata as populateMetaData
 * (Cullst/loadCullData.c:95-118). */
int st_write_cull(const char *outBase, const char *inBase, const st_cull_meta *meta, const float *offRS,
                  const float *offAS, const float *sigmaR, const float *sigmaA, const float *corr,
                  const uint8_t *type, int32_t nR, int32_t nA, int32_t byteOrder);

/* What the Cullls header records (matchResult / lsFit fields of the
 * un-vendored landsatSource64). */
typedef struct st_cull_ls_meta {
    const char *fileEarly, *fileLate;
    double x0, y0, dx, dy;
    uint32_t stepX, stepY;
    int32_t slowFlag, proj;
    double jdEarly, jdLate;
    float noData; /* NODATA: a point counts as good when XS and YS exceed NODATA + 1 */
} st_cull_ls_meta;

/* writeLSCulledOffsets (Cullls/cullls.c:63-175): matchFile.cull.dx .dy .sx
 * .sy (float32, byteOrder; the reference writes MSB) and .cull.mtype, and
 * matchFile.cull.dat with the reference's printf formats, the success and
 * culled rates from nMatch / nAttempt (st_cull_ls_params) and the RMS
 * sigmas over the good points. */
int st_write_cull_ls(const char *matchFile, const st_cull_ls_meta *meta, const float *XS, const float *YS,
                     const float *sigmaX, const float *sigmaY, const uint8_t *type, int32_t nx, int32_t ny,
                     int32_t nMatch, int32_t nAttempt, int32_t byteOrder);

/* ---- inputs before tracking ------------------------------------------ */

/* parseInitialOffsets (Strack/parseInitialOffsets.c:11-64): the
 * "range_offset_polynomial:" and "azimuth_offset_polynomial:" lines of an
 * initial-offset file, three coefficients each (st_track_params
 * rShiftPoly / aShiftPoly with polyShift = 1). */
int st_read_offset_polys(const char *file, double rShiftPoly[3], double aShiftPoly[3]);

/* readBothOffsetsStrack (Strack/readBothOffsetsStrack.c:76-120): the header
 * of daFile.dat ("%i%i%i%i%f%f": r0 a0 nr na deltaR deltaA) ... */
int st_read_initial_offsets_header(const char *daFile, int32_t *r0, int32_t *a0, int32_t *nr, int32_t *na,
                                   float *deltaR, float *deltaA);
/* ... and the two grids (na x nr big-endian float32: azimuth offsets from
 * daFile, range offsets from drFile) for st_track_params initDa / initDr. */
int st_read_initial_offsets(const char *daFile, const char *drFile, int32_t nr, int32_t na, float *da, float *dr);

/* getMask, offset-mask branch (Strack/getMask.c:75-118): the header of the
 * mask's .dat file into hdr[6] = r0 a0 nr na nrl nal (nrl, nal = the
 * integer parts of deltaR, deltaA) ... */
int st_read_mask_header(const char *datFile, int32_t hdr[6]);
/* ... and the na x nr mask bytes (st_track_params mask). */
int st_read_mask(const char *maskFile, int32_t nr, int32_t na, uint8_t *mask);

/* getInt (Strack/getInt.c:11-77): the complex interferogram, big-endian
 * int16 (re, im) pairs, na lines of nr samples, into intf (2 * nr * na
 * floats, st_track_params intf).  A file with fewer lines leaves the rest
 * at 1 + 1i, as the reference's fill does; nr, na and the looks come from
 * the interferogram's geodat file, which the caller reads. */
int st_read_interferogram(const char *intFile, int32_t nr, int32_t na, float *intf);

/* parseTrack (Strack/parseTrack.c:13-190): the strack parameter file, an
 * RDF keyword file with one "keyword (units) = value ! comment" per line
 * (units and comment optional).  The reference reads it through GrIMP's
 * rdfSource, which is not in the reference tree, so the line rules here
 * are parity unpinned: keywords compare without regard to case, the first
 * occurrence of a keyword wins, and blank or comment-only lines are
 * skipped.  The keywords, their %i formats, which are required, the
 * defaults and the quirks follow parseTrack.c:
 *   - edgeR / edgeA override edgepad; an unreadable edgeA resets edgePadR
 *     (parseTrack.c:155-159);
 *   - rStart, aStart, deltaR, deltaA are divided by scalefactor (C integer
 *     division, parseTrack.c:165-168);
 *   - offsetmaskfile without offsetmaskdat leaves maskGeodat empty
 *     (parseTrack.c:73-77);
 *   - intfile and baseparams each require intgeodat; noComplex (the
 *     command-line flag) clears baseFlag (parseTrack.c:89-109).
 * An image with a VRT beside it ("<image>.vrt", checkForVrt in the
 * un-vendored clib: parity unpinned) needs no image1par / image2par.
 * Missing required keywords give ST_ERR_ARG with parseTrack's message. */
#define ST_PAR_PATH 1024
#define ST_GEODATMASK 0 /* strack.h:5 */
#define ST_OFFSETMASK 1 /* strack.h:6 */
typedef struct st_track_par_file {
    char image1[ST_PAR_PATH], image2[ST_PAR_PATH];
    char par1[ST_PAR_PATH], par2[ST_PAR_PATH]; /* image1par / image2par ("" with a VRT) */
    char vrt1[ST_PAR_PATH], vrt2[ST_PAR_PATH]; /* "" without a VRT */
    char initShift[ST_PAR_PATH];               /* initial-offset grids (polyShift = 0) */
    char initialOffsetFile[ST_PAR_PATH];       /* offset polynomials (polyShift = 1) */
    char outputFile[ST_PAR_PATH];              /* base of the .dr .da .cc .mt .dat .azd .vrt .mt.vrt products */
    char maskFile[ST_PAR_PATH], maskGeodat[ST_PAR_PATH], maskVrt[ST_PAR_PATH];
    char intFile[ST_PAR_PATH], intGeodat[ST_PAR_PATH], baseParams[ST_PAR_PATH];
    int32_t polyShift, maskFlag, maskType, intFlag, baseFlag;
    int32_t rStart, aStart, deltaR, deltaA, nR, nA, wR, wA, wRa, wAa;
    int32_t edgePad, edgePadR, edgePadA, scaleFactor, navgR, navgA;
} st_track_par_file;
int st_read_track_par(const char *parFile, int32_t noComplex, st_track_par_file *out);

/* ---- GDAL VRT inputs (raw bands) -------------------------------------- */
/* The VRT files GrIMP reads through GDAL (GDALOpen / GDALRasterIO /
 * readDataSetMetaData) where the bands are VRTRawRasterBand, as its own
 * writers produce them: readBothOffsetsStrackVrt
 * (Strack/readBothOffsetsStrack.c:46-76), getMaskVRT (Strack/getMask.c:22-48),
 * loadFromVRT (Cullst/loadCullData.c:46-92).  Other band kinds return
 * ST_ERR_UNSUPPORTED.  Errors: st_vrt_last_error(). */
typedef struct st_vrt st_vrt;

const char *st_vrt_last_error(void);
int st_vrt_open(const char *vrtFile, st_vrt **vrt);
void st_vrt_close(st_vrt *vrt);
/* raster size (GDALGetRasterXSize / YSize) and band count */
int st_vrt_size(const st_vrt *vrt, int32_t *nx, int32_t *ny, int32_t *nBands);
/* a default-domain dataset metadata item (readDataSetMetaData + get_value) */
int st_vrt_metadata(const st_vrt *vrt, const char *key, char *value, size_t len);
/* band (1-based) description, or its "Description" metadata item */
int st_vrt_band_description(const st_vrt *vrt, int32_t band, char *desc, size_t len);
/* band converted to float32 (GDALRasterIO with GDT_Float32), ny x nx */
int st_vrt_read_f32(const st_vrt *vrt, int32_t band, float *out);
/* Byte band (GDT_Byte) */
int st_vrt_read_u8(const st_vrt *vrt, int32_t band, uint8_t *out);
/* CInt16 / CFloat32 band as (re, im) float pairs (SLC samples) */
int st_vrt_read_c64(const st_vrt *vrt, int32_t band, float *out);

#ifdef __cplusplus
}
#endif

#endif /* SPECKLE_IO_H */
