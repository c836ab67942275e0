// ozl_errors.h -- library error codes (no HIP dependency).
#pragma once

// Error codes: the numeric values of the reference's ZL_ErrorCode
// (include/openzl/zl_errors_types.h:41-99) so callers see the same codes.
enum {
    OZL_OK = 0,
    OZL_ERR_GENERIC = 1,
    OZL_ERR_SRCSIZE_TOOSMALL = 3,
    OZL_ERR_DSTCAPACITY_TOOSMALL = 5,
    OZL_ERR_HEADER_UNKNOWN = 10,
    OZL_ERR_CORRUPTION = 12,
    OZL_ERR_COMPRESSED_CHECKSUM = 13,
    OZL_ERR_CONTENT_CHECKSUM = 14,
    OZL_ERR_PARAMETER_INVALID = 21,
    OZL_ERR_GRAPH_INVALID = 30,
    OZL_ERR_INPUT_TYPE_UNSUPPORTED = 43,
    OZL_ERR_SUCCESSOR_INVALID = 40,
    OZL_ERR_NODE_INVALID_INPUT = 55,
    OZL_ERR_FORMAT_VERSION_UNSUPPORTED = 60,
    OZL_ERR_ALLOCATION = 70,
    OZL_ERR_INTERNAL_BUFFER = 71,
    OZL_ERR_LOGIC = 80,
    OZL_ERR_HIP = 90,  // a HIP runtime call failed (no reference counterpart)
};
