/*
 * Test harness: a small stand-in for MATLAB's MEX API (R2018a interleaved
 * complex), enough to compile and run the MEX shims of
 * faceted-hypersara_amd/mex/ outside MATLAB.  Only what the shims use:
 * double (real / complex, dense / sparse), uint64, char, cell and 1x1
 * struct arrays.  Implemented in mexmock.c; mexErrMsgIdAndTxt longjmps back
 * to mexmock_call, which frees the call's mxMalloc temporaries as MATLAB
 * does.  Test infrastructure only.
 */
#ifndef FHS_MEXMOCK_MEX_H
#define FHS_MEXMOCK_MEX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;
typedef struct {
  double real, imag;
} mxComplexDouble;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef enum { mxUNKNOWN_CLASS = 0, mxCELL_CLASS, mxSTRUCT_CLASS, mxCHAR_CLASS, mxDOUBLE_CLASS,
               mxUINT64_CLASS } mxClassID;

void* mxMalloc(size_t n);
void mxFree(void* p);

size_t mxGetNumberOfElements(const mxArray* a);
size_t mxGetM(const mxArray* a);
size_t mxGetN(const mxArray* a);
void mxSetM(mxArray* a, size_t m);
void mxSetN(mxArray* a, size_t n);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
int mxIsCell(const mxArray* a);
int mxIsStruct(const mxArray* a);
int mxIsComplex(const mxArray* a);
int mxIsSparse(const mxArray* a);
int mxIsUint64(const mxArray* a);
double mxGetScalar(const mxArray* a);
double* mxGetDoubles(const mxArray* a);
mxComplexDouble* mxGetComplexDoubles(const mxArray* a);
void* mxGetData(const mxArray* a);
mwIndex* mxGetIr(const mxArray* a);
mwIndex* mxGetJc(const mxArray* a);
mxArray* mxGetCell(const mxArray* a, mwIndex i);
void mxSetCell(mxArray* a, mwIndex i, mxArray* v);
mxArray* mxGetField(const mxArray* a, mwIndex i, const char* name);
void mxSetField(mxArray* a, mwIndex i, const char* name, mxArray* v);
int mxGetString(const mxArray* a, char* buf, mwSize len);

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity c);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity c);
mxArray* mxCreateCellMatrix(mwSize m, mwSize n);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names);
mxArray* mxCreateString(const char* s);
mxArray* mxCreateSparse(mwSize m, mwSize n, mwSize nzmax, mxComplexity c);
mxArray* mxDuplicateArray(const mxArray* a);
void mxDestroyArray(mxArray* a);

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
int mexAtExit(void (*fn)(void));
void mexLock(void);
int mexCallMATLAB(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name);

#ifdef __cplusplus
}
#endif
#endif
