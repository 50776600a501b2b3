// aenet_capi_status.h -- status codes of the C ABI (values of the
// reference's AENET_* globals, src/aenet.f90:101-105) plus this library's
// own codes above them.
#pragma once
#define AENET_STATUS_OK 0
#define AENET_STATUS_ERR_INIT 1
#define AENET_STATUS_ERR_MALLOC 2
#define AENET_STATUS_ERR_IO 3
#define AENET_STATUS_ERR_TYPE 4
#define AENET_STATUS_ERR_CAPACITY 5   // a neighbor/cell capacity was exceeded
#define AENET_STATUS_ERR_DEVICE 6     // HIP runtime error or no usable GPU
#define AENET_STATUS_ERR_ARG 7        // invalid argument
#define AENET_STATUS_ERR_COORD 8      // non-finite coordinates, or an isolated structure left its bounding box
#define AENET_AMD_ERR_ARG_VALUE AENET_STATUS_ERR_ARG
