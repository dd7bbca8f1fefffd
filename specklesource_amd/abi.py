"""ctypes mirror of include/speckle_amd.h and include/speckle_synth.h.

The structures here must match the C headers field for field; the
layout test in tests/test_abi.py checks the sizes against the compiled
library.
"""
import ctypes as C

ST_OK = 0
ST_ERR_ARG = -1
ST_ERR_HIP = -2
ST_ERR_UNSUPPORTED = -3
ST_ERR_NOMEM = -4
ST_ERR_NODEV = -5
ST_ERR_CANCELLED = -7
ST_KERNEL_SLOTS = 16  # kernel timing slots of st_track_stats

ST_SAMPLE_C64 = 0
ST_SAMPLE_CI16 = 1

ST_BAD, ST_CMATCH, ST_AMPMATCH, ST_AMPMATCHLARGE = 0, 1, 2, 3
ST_NODATA = -2.0e9

ST_PREC_FP32 = 0
ST_PREC_MFMA = 1
ST_PREC_AUTO = 2

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int32)
_u8 = C.POINTER(C.c_uint8)


class TrackParams(C.Structure):
    """st_track_params: TrackParams fields of Strack/strack.h:56-165."""

    _fields_ = [
        ("rStart", C.c_int32), ("aStart", C.c_int32),
        ("deltaR", C.c_int32), ("deltaA", C.c_int32),
        ("nR", C.c_int32), ("nA", C.c_int32),
        ("wR", C.c_int32), ("wA", C.c_int32),
        ("wRa", C.c_int32), ("wAa", C.c_int32),
        ("edgePadR", C.c_int32), ("edgePadA", C.c_int32),
        ("navgR", C.c_int32), ("navgA", C.c_int32),
        ("scaleFactor", C.c_int32),
        ("noComplex", C.c_int32), ("hanningFlag", C.c_int32),
        ("gaussFlag", C.c_int32), ("maxTries", C.c_int32),
        ("checkAzFocus", C.c_int32),
        ("azDefocusThresh", C.c_float),
        ("precision", C.c_int32),
        ("polyShift", C.c_int32),
        ("rShiftPoly", C.c_double * 3), ("aShiftPoly", C.c_double * 3),
        ("initDr", _f), ("initDa", _f),
        ("initNr", C.c_int32), ("initNa", C.c_int32),
        ("initR0", C.c_double), ("initA0", C.c_double),
        ("initDeltaR", C.c_double), ("initDeltaA", C.c_double),
        ("mask", _u8),
        ("maskNr", C.c_int32), ("maskNa", C.c_int32),
        ("maskR0", C.c_int32), ("maskA0", C.c_int32),
        ("maskNrl", C.c_int32), ("maskNal", C.c_int32),
        ("baseFlag", C.c_int32),
        ("Bn", C.c_double), ("Bp", C.c_double), ("dBn", C.c_double),
        ("dBp", C.c_double), ("dBnQ", C.c_double), ("dBpQ", C.c_double),
        ("lambda_", C.c_double),
        ("nSlpA1", C.c_int32),
        ("sinThetaD", _d), ("cosThetaD", _d), ("overRange", _d),
        ("intf", _f),
        ("intNr", C.c_int32), ("intNa", C.c_int32), ("intNrl", C.c_int32), ("intNal", C.c_int32),
    ]


class TrackOut(C.Structure):
    _fields_ = [
        ("offR", _f), ("offA", _f), ("corr", _f), ("azd", _f), ("type", _u8),
        ("nComplex", C.c_int32), ("nAmp", C.c_int32),
        ("nAmpL", C.c_int32), ("nFail", C.c_int32),
    ]


class TrackDiag(C.Structure):
    _fields_ = [
        ("r2", _i), ("a2", _i), ("azShift", _i),
        ("peakI", _i * 3), ("peakJ", _i * 3),
        ("fineI", _i * 3), ("fineJ", _i * 3),
        ("cmax", _d * 3),
        ("sigRg", _f * 3), ("sigAz", _f * 3),
        ("peakGap", _f * 3), ("fineGap", _f * 3),
        ("snr", _f * 3),
    ]


class TrackStats(C.Structure):
    _fields_ = [
        ("nPoints", C.c_int64), ("nCorr", C.c_int64 * 3), ("stageMs", C.c_double * 4),
        ("totalMs", C.c_double), ("kernelMs", C.c_double * ST_KERNEL_SLOTS),
        ("kernelLaunches", C.c_int64 * ST_KERNEL_SLOTS), ("kernelNames", (C.c_char * 32) * ST_KERNEL_SLOTS),
    ]


class CullParams(C.Structure):
    """st_cull_params: CullParams fields of Cullst/cullst.h:3-60."""

    _fields_ = [
        ("nR", C.c_int32), ("nA", C.c_int32),
        ("offR", _f), ("offA", _f), ("corr", _f), ("type", _u8),
        ("mask", _u8),
        ("offRS", _f), ("offAS", _f), ("sigmaR", _f), ("sigmaA", _f),
        ("bR", C.c_int32), ("bA", C.c_int32), ("nGood", C.c_int32),
        ("maxA", C.c_float), ("maxR", C.c_float),
        ("sR", C.c_int32), ("sA", C.c_int32),
        ("islandThresh", C.c_int32), ("singleMT", C.c_int32),
        ("corrThresh", C.c_float), ("nCulled", C.c_int32),
    ]


class CullLSParams(C.Structure):
    """st_cull_ls_params: CullLSParams fields of Cullls/cullls.h:2-28 plus
    the matchResult / lsFit fields the stages read."""

    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("X", _f), ("Y", _f), ("type", _u8),
        ("XS", _f), ("YS", _f), ("sigmaX", _f), ("sigmaY", _f),
        ("bR", C.c_int32), ("bA", C.c_int32), ("nGood", C.c_int32),
        ("maxX", C.c_float), ("maxY", C.c_float),
        ("sX", C.c_int32), ("sY", C.c_int32), ("islandThresh", C.c_int32),
        ("deltaT", C.c_double), ("dx", C.c_double), ("dy", C.c_double),
        ("slowVelMask", C.c_uint8), ("noDataToUseMask", C.c_uint8),
        ("noData", C.c_float), ("nover", C.c_int32),
        ("nMatch", C.c_int32), ("nAttempt", C.c_int32), ("nCulled", C.c_int32),
    ]


ST_JOB_STRACK, ST_JOB_STRACKW, ST_JOB_CULL = 0, 1, 2
ST_JOB_QUEUED, ST_JOB_RUNNING, ST_JOB_DONE, ST_JOB_FAILED, ST_JOB_CANCELLED = 0, 1, 2, 3, 4


class Job(C.Structure):
    """st_job: one tracking request of a server."""

    _fields_ = [
        ("kind", C.c_int32), ("pair", C.c_int32),
        ("img1", C.c_void_p), ("img2", C.c_void_p),
        ("nr1", C.c_int32), ("na1", C.c_int32), ("nr2", C.c_int32), ("na2", C.c_int32),
        ("sampleType", C.c_int32),
        ("par", TrackParams), ("out", TrackOut),
        ("cull", C.POINTER(CullParams)),
    ]


class JobInfo(C.Structure):
    _fields_ = [
        ("state", C.c_int32), ("status", C.c_int32), ("slot", C.c_int32), ("pad", C.c_int32),
        ("queueMs", C.c_double), ("runMs", C.c_double), ("latencyMs", C.c_double),
        ("patches", C.c_int64),
        ("nComplex", C.c_int32), ("nAmp", C.c_int32), ("nAmpL", C.c_int32), ("nFail", C.c_int32),
        ("error", C.c_char * 128),
    ]


class SynthSpec(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("na", C.c_int32), ("seed", C.c_uint64),
        ("dr0", C.c_double), ("drR", C.c_double), ("drA", C.c_double),
        ("da0", C.c_double), ("daR", C.c_double), ("daA", C.c_double),
        ("ridgeR", C.c_double), ("ridgeA", C.c_double),
        ("ridgeC", C.c_double), ("ridgeW", C.c_double),
        ("shearR", C.c_double), ("shearA", C.c_double),
        ("gmin", C.c_double), ("gmax", C.c_double),
        ("gLr", C.c_double), ("gLa", C.c_double),
        ("sigmaLn", C.c_double), ("corrLen", C.c_double),
        ("bwR", C.c_double), ("bwA", C.c_double), ("fdc", C.c_double),
        ("maskR0", C.c_int32), ("maskR1", C.c_int32),
        ("maskA0", C.c_int32), ("maskA1", C.c_int32),
    ]


def ptr(arr, ctype):
    """ctypes pointer to a numpy array's data (None passes NULL)."""
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))
