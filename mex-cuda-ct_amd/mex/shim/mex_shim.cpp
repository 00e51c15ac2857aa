// mex_shim.cpp -- implementation of the minimal MEX/matrix API of
// shim/mex.h: owned numeric arrays, scalar reads and exceptions for the
// error calls.  Used only to test the gateways without MATLAB.
#include "mex.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <exception>
#include <string>
#include <vector>

struct mxArray_tag {
  mxClassID cls = mxUNKNOWN_CLASS;
  std::vector<mwSize> dims;
  std::vector<unsigned char> data;
};

namespace {

struct mexct_shim_error : std::exception {
  std::string id, msg;
  mexct_shim_error(const char* i, const char* m) : id(i ? i : ""), msg(m ? m : "") {}
  const char* what() const noexcept override { return msg.c_str(); }
};

thread_local std::string g_id, g_msg;

size_t elem_size(mxClassID c) {
  switch (c) {
    case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
    case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
    case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
    case mxINT8_CLASS: case mxUINT8_CLASS: case mxLOGICAL_CLASS: return 1;
    default: return 0;
  }
}

size_t numel(const mxArray* a) {
  size_t n = 1;
  for (mwSize d : a->dims) n *= d;
  return n;
}

}  // namespace

extern "C" {

void* mxGetData(const mxArray* pa) { return pa ? (void*)pa->data.data() : nullptr; }
const mwSize* mxGetDimensions(const mxArray* pa) { return pa->dims.data(); }
mwSize mxGetNumberOfDimensions(const mxArray* pa) { return pa->dims.size(); }
size_t mxGetNumberOfElements(const mxArray* pa) { return numel(pa); }
mxClassID mxGetClassID(const mxArray* pa) { return pa->cls; }

double mxGetScalar(const mxArray* pa) {
  if (!pa || numel(pa) == 0) return 0.0;
  const void* p = pa->data.data();
  switch (pa->cls) {
    case mxDOUBLE_CLASS: return *(const double*)p;
    case mxSINGLE_CLASS: return *(const float*)p;
    case mxINT32_CLASS: return *(const int32_t*)p;
    case mxUINT32_CLASS: return *(const uint32_t*)p;
    case mxINT64_CLASS: return (double)*(const int64_t*)p;
    case mxUINT64_CLASS: return (double)*(const uint64_t*)p;
    case mxINT16_CLASS: return *(const int16_t*)p;
    case mxUINT16_CLASS: return *(const uint16_t*)p;
    case mxINT8_CLASS: return *(const int8_t*)p;
    case mxUINT8_CLASS: case mxLOGICAL_CLASS: return *(const uint8_t*)p;
    default: return 0.0;
  }
}

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid,
                              mxComplexity flag) {
  (void)flag;
  mxArray* a = new mxArray_tag();
  a->cls = classid;
  a->dims.assign(dims, dims + ndim);
  while (a->dims.size() > 2 && a->dims.back() == 1) a->dims.pop_back();  // MATLAB drops trailing 1s
  if (a->dims.size() < 2) a->dims.resize(2, 1);
  a->data.assign(numel(a) * elem_size(classid), 0);  // zero-filled, as MATLAB's
  return a;
}

void mexErrMsgIdAndTxt(const char* errorid, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw mexct_shim_error(errorid, buf);
}

void mexErrMsgTxt(const char* msg) { throw mexct_shim_error("", msg); }

mxArray* shim_new(int classid, int ndim, const size_t* dims, const void* data) {
  std::vector<mwSize> d(dims, dims + ndim);
  mxArray* a = new mxArray_tag();
  a->cls = (mxClassID)classid;
  a->dims = d;
  if (a->dims.size() < 2) a->dims.resize(2, 1);
  a->data.resize(numel(a) * elem_size(a->cls));
  if (data && !a->data.empty()) memcpy(a->data.data(), data, a->data.size());
  return a;
}

void shim_free(mxArray* pa) { delete pa; }

int shim_invoke(shim_mex_fn fn, int nlhs, mxArray** plhs, int nrhs, const mxArray** prhs) {
  g_id.clear();
  g_msg.clear();
  try {
    fn(nlhs, plhs, nrhs, prhs);
  } catch (const mexct_shim_error& e) {
    g_id = e.id;
    g_msg = e.msg;
    return -1;
  }
  return 0;
}

const char* shim_last_id(void) { return g_id.c_str(); }
const char* shim_last_msg(void) { return g_msg.c_str(); }

}  // extern "C"
