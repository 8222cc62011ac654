// Host-memory implementation of the MEX API test double declared in mex.h.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

struct mxArray_tag {
  mxClassID cls = mxUNKNOWN_CLASS;
  size_t m = 0, n = 0;
  std::vector<unsigned char> data;       // numeric / logical payload, column-major
  bool sparse = false;
  std::vector<mwIndex> ir, jc;           // sparse: row indices, column starts
  std::vector<std::string> fields;       // struct: field names
  std::vector<mxArray*> slots;           // struct: numel x fields; cell: numel
  std::map<std::string, mxArray*> props; // object properties
};

namespace {
long g_live = 0;

size_t elem_size(mxClassID c) {
  switch (c) {
    case mxLOGICAL_CLASS: return sizeof(mxLogical);
    case mxINT8_CLASS: case mxUINT8_CLASS: return 1;
    case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
    case mxINT32_CLASS: case mxUINT32_CLASS: case mxSINGLE_CLASS: return 4;
    case mxINT64_CLASS: case mxUINT64_CLASS: case mxDOUBLE_CLASS: return 8;
    default: return 0;
  }
}

mxArray* make(mxClassID c, size_t m, size_t n) {
  mxArray* a = new mxArray_tag;
  ++g_live;
  a->cls = c;
  a->m = m;
  a->n = n;
  a->data.assign(m * n * elem_size(c), 0);
  return a;
}

int field_index(const mxArray* a, const char* name) {
  for (size_t f = 0; f < a->fields.size(); ++f)
    if (a->fields[f] == name) return (int)f;
  return -1;
}
}  // namespace

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity) { return make(cls, m, n); }
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity) { return make(mxDOUBLE_CLASS, m, n); }
mxArray* mxCreateLogicalScalar(mxLogical v) {
  mxArray* a = make(mxLOGICAL_CLASS, 1, 1);
  *(mxLogical*)a->data.data() = v;
  return a;
}
mxArray* mxCreateSparseLogicalMatrix(mwSize m, mwSize n, mwSize nzmax) {
  mxArray* a = make(mxLOGICAL_CLASS, 0, 0);
  a->m = m;
  a->n = n;
  a->sparse = true;
  a->data.assign(nzmax * sizeof(mxLogical), 0);
  a->ir.assign(nzmax, 0);
  a->jc.assign(n + 1, 0);
  return a;
}
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names) {
  mxArray* a = make(mxSTRUCT_CLASS, m, n);
  for (int f = 0; f < nfields; ++f) a->fields.push_back(names[f]);
  a->slots.assign(m * n * nfields, nullptr);
  return a;
}
mxArray* mxCreateCellMatrix(mwSize m, mwSize n) {
  mxArray* a = make(mxCELL_CLASS, m, n);
  a->slots.assign(m * n, nullptr);
  return a;
}
void mxDestroyArray(mxArray* a) {
  if (!a) return;
  for (mxArray* s : a->slots) mxDestroyArray(s);
  for (auto& kv : a->props) mxDestroyArray(kv.second);
  delete a;
  --g_live;
}

void* mxGetData(const mxArray* a) { return a->data.empty() ? nullptr : (void*)a->data.data(); }
double* mxGetPr(const mxArray* a) { return a->cls == mxDOUBLE_CLASS ? (double*)mxGetData(a) : nullptr; }
mwIndex* mxGetIr(const mxArray* a) { return a->sparse ? (mwIndex*)a->ir.data() : nullptr; }
mwIndex* mxGetJc(const mxArray* a) { return a->sparse ? (mwIndex*)a->jc.data() : nullptr; }
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
size_t mxGetM(const mxArray* a) { return a->m; }
size_t mxGetN(const mxArray* a) { return a->n; }
size_t mxGetNumberOfElements(const mxArray* a) { return a->m * a->n; }
bool mxIsComplex(const mxArray*) { return false; }
bool mxIsSparse(const mxArray* a) { return a->sparse; }
bool mxIsCell(const mxArray* a) { return a->cls == mxCELL_CLASS; }
bool mxIsStruct(const mxArray* a) { return a->cls == mxSTRUCT_CLASS; }
mxArray* mxGetCell(const mxArray* a, mwIndex i) { return i < a->slots.size() ? a->slots[i] : nullptr; }
void mxSetCell(mxArray* a, mwIndex i, mxArray* v) {
  mxDestroyArray(a->slots.at(i));
  a->slots.at(i) = v;
}
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name) {
  const int f = a->cls == mxSTRUCT_CLASS ? field_index(a, name) : -1;
  if (f < 0 || i >= a->m * a->n) return nullptr;
  return a->slots[i * a->fields.size() + f];
}
void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* v) {
  const int f = field_index(a, name);
  if (f < 0) throw std::runtime_error(std::string("mexsim: no field ") + name);
  mxArray*& slot = a->slots.at(i * a->fields.size() + f);
  mxDestroyArray(slot);
  slot = v;
}
mxArray* mxGetProperty(const mxArray* a, mwIndex, const char* name) {
  auto it = a->props.find(name);
  return it == a->props.end() ? nullptr : it->second;
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw mexsim::Error(id, buf);
}
int mexLock(void) { return 0; }
int mexUnlock(void) { return 0; }

namespace mexsim {
mxArray* object_with_property(const char* name, mxArray* value) {
  mxArray* a = make(mxUNKNOWN_CLASS, 1, 1);
  a->props[name] = value;
  return a;
}
long live_arrays() { return g_live; }
}  // namespace mexsim
