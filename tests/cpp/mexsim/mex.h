// Test double of the MATLAB MEX C API: the subset fecl_amd/matlab/fecl_wrap.cpp uses, with the
// documented names and signatures, so the gateway can be built and driven without MATLAB.
// Arrays live in host memory; mexErrMsgIdAndTxt throws mexsim::Error instead of returning to
// the MATLAB prompt.  Not used by the library itself.
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef bool mxLogical;
typedef char mxChar;

typedef enum {
  mxUNKNOWN_CLASS,
  mxCELL_CLASS,
  mxSTRUCT_CLASS,
  mxLOGICAL_CLASS,
  mxCHAR_CLASS,
  mxVOID_CLASS,
  mxDOUBLE_CLASS,
  mxSINGLE_CLASS,
  mxINT8_CLASS,
  mxUINT8_CLASS,
  mxINT16_CLASS,
  mxUINT16_CLASS,
  mxINT32_CLASS,
  mxUINT32_CLASS,
  mxINT64_CLASS,
  mxUINT64_CLASS,
  mxFUNCTION_CLASS
} mxClassID;

typedef enum { mxREAL, mxCOMPLEX } mxComplexity;

struct mxArray_tag;
typedef struct mxArray_tag mxArray;

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity flag);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray* mxCreateLogicalScalar(mxLogical value);
mxArray* mxCreateSparseLogicalMatrix(mwSize m, mwSize n, mwSize nzmax);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** fieldnames);
mxArray* mxCreateCellMatrix(mwSize m, mwSize n);
void mxDestroyArray(mxArray* a);

void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
mwIndex* mxGetIr(const mxArray* a);
mwIndex* mxGetJc(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
bool mxIsComplex(const mxArray* a);
bool mxIsSparse(const mxArray* a);
bool mxIsCell(const mxArray* a);
bool mxIsStruct(const mxArray* a);
mxArray* mxGetCell(const mxArray* a, mwIndex i);
void mxSetCell(mxArray* a, mwIndex i, mxArray* value);
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name);
void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* value);
mxArray* mxGetProperty(const mxArray* a, mwIndex i, const char* name);

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexLock(void);
int mexUnlock(void);

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);

// ---- test-double extensions (not part of the MEX API)
namespace mexsim {
struct Error : std::runtime_error {
  std::string id;
  Error(const std::string& i, const std::string& m) : std::runtime_error(m), id(i) {}
};
// A MATLAB object whose property `name` holds `value` (FeCl's codec objects keep their
// handle in the property mexHandle_).
mxArray* object_with_property(const char* name, mxArray* value);
// Arrays alive (created and not destroyed), for leak checks.
long live_arrays();
}  // namespace mexsim
