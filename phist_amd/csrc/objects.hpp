// objects.hpp -- typed implementations behind the C ABI (exports.cpp).
#pragma once
#include "amd_common.hpp"

namespace phx {

template <typename T>
SparseMat<T>* sparse_from_rowfunc(const Comm* comm, int64_t nrows, int64_t ncols, int32_t maxnne,
                                  phist_sparseMat_rowFunc f, phist_sparseMat_rowFuncConstructor ctor, void* arg,
                                  bool parallel);
template <typename T>
SparseMat<T>* sparse_from_mm(const Comm* comm, const char* filename);

namespace api {

template <typename T> Mvec<T>* mvec_create(const Map* map, int nvec);
template <typename T> void mvec_delete(Mvec<T>* V);
template <typename T> void mvec_view_block(Mvec<T>* V, Mvec<T>** Vblock, int jmin, int jmax);
template <typename T> void mvec_get_block(const Mvec<T>* V, Mvec<T>* B, int jmin, int jmax);
template <typename T> void mvec_set_block(Mvec<T>* V, const Mvec<T>* B, int jmin, int jmax);
template <typename T> void mvec_to_device(Mvec<T>* V);
template <typename T> void mvec_from_device(Mvec<T>* V);
template <typename T> void mvec_get_data(const Mvec<T>* V, T* data, int64_t lda, int row_major);
template <typename T> void mvec_set_data(Mvec<T>* V, const T* data, int64_t lda, int row_major);
template <typename T> void mvec_put_func(Mvec<T>* V, phist_mvec_elemFunc f, void* arg);
template <typename T> void mvec_random(Mvec<T>* V);
template <typename T> void mvec_print(const Mvec<T>* V);
template <typename T> void mvec_write_bin(const Mvec<T>* V, const char* fn);
template <typename T> void mvec_read_bin(Mvec<T>* V, const char* fn);

template <typename T> SdMat<T>* sdMat_create(int nrows, int ncols);
template <typename T> void sdMat_delete(SdMat<T>* M);
template <typename T> void sdMat_view_block(SdMat<T>* M, SdMat<T>** Mb, int imin, int imax, int jmin, int jmax);
template <typename T> void sdMat_get_block(const SdMat<T>* M, SdMat<T>* B, int imin, int imax, int jmin, int jmax);
template <typename T> void sdMat_set_block(SdMat<T>* M, const SdMat<T>* B, int imin, int imax, int jmin, int jmax);
template <typename T> void sdMat_set_data(SdMat<T>* M, const T* data, int64_t lda, int row_major);
template <typename T> void sdMat_get_data(const SdMat<T>* M, T* data, int64_t lda, int row_major);
template <typename T> void sdMat_put_value(SdMat<T>* M, T v);
template <typename T> void sdMat_identity(SdMat<T>* M);
template <typename T> void sdMat_random(SdMat<T>* M);
template <typename T> void sdMat_print(const SdMat<T>* M);
template <typename T> void sdMat_write_bin(const SdMat<T>* M, const char* fn);
template <typename T> void sdMat_read_bin(SdMat<T>* M, const char* fn);
template <typename T>
void sdMat_gemm(T alpha, const SdMat<T>* A, bool ta, const SdMat<T>* B, bool tb, T beta, SdMat<T>* C);
template <typename T> void sdMat_add(T alpha, const SdMat<T>* A, bool ta, T beta, SdMat<T>* B);

template <typename T> void sdMat_cholesky(SdMat<T>* M, int* perm, int* rank, T rankTol);
template <typename T> void sdMat_subst(const SdMat<T>* R, const int* perm, int rank, SdMat<T>* X, bool backward);
template <typename T> void sdMat_qb(SdMat<T>* B, SdMat<T>* B_1, T* nrms, int* rank, T rankTol);
template <typename T> int chol_QRp(Mvec<T>* V, SdMat<T>* R, int* perm, bool robust);
template <typename T> int svqb(Mvec<T>* V, SdMat<T>* B, T* e, bool robust);
template <typename T>
int orthog(const Mvec<T>* V, Mvec<T>* W, SdMat<T>* R1, SdMat<T>* R2, int numSweeps, int* rankVW, int flags);

}  // namespace api
}  // namespace phx
