/*
 * oracle.h -- CPU restatement of tomograph/synkrotomo's projector and SIRT (test
 * infrastructure only; see oracle.c).
 */
#ifndef SYNKRO_ORACLE_H
#define SYNKRO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int orc_preprocess(const float* angles, int A, float* cost, float* sint, float* lbase, int* orig);

void orc_fp(const float* angles, int A, float rho0, float drho, int R, const float* img, int n,
            float* out, int nt);
void orc_bp(const float* angles, int A, float rho0, float drho, int size, const float* proj,
            int p, float* out, int nt);
void orc_sirt(const float* angles, int A, float rho0, float drho, float* img, int n,
              const float* proj, int p, int iters, int nt);
void orc_sirt3d(const float* angles, int A, const float* rhos, int r, const float* proj,
                int iters, int size, float* out, int nt);

void orc_fp_f64(const float* angles, int A, double rho0, double drho, int R, const double* img,
                int n, double* out, int nt);
void orc_bp_f64(const float* angles, int A, double rho0, double drho, int size,
                const double* proj, int p, double* out, int nt);
void orc_sirt_f64(const float* angles, int A, double rho0, double drho, double* img, int n,
                  const double* proj, int p, int iters, int nt);

float orc_fp_ray_line(float cost, float sint, float lbase, int r, float rho0, float drho, int h,
                      const float* frame_img);
float orc_bp_pixel_line(float cost, float sint, float lbase, float xmin, float ymin, float rho0,
                        float drho, int ncand, int R, const float* proj_row, int* s_out,
                        float* l_out);

int64_t orc_fp_csr(const float* angles, int A, float rho0, float drho, int R, int N,
                   int64_t* rowptr, int32_t* cols, float* vals);
int64_t orc_bp_csr(const float* angles, int A, float rho0, float drho, int R, int N,
                   int64_t* rowptr, int32_t* cols, float* vals);
void orc_fp_adjoint(const float* angles, int A, float rho0, float drho, int R, int N,
                    const float* sino, double* out);
int64_t orc_bp_oob_nonzero(const float* angles, int A, float rho0, float drho, int R, int N);

int orc_sanity_check(const float* sane, const float* test, int64_t n, float tol);
int64_t orc_diff(const float* sane, const float* test, int64_t n, float tol);
float orc_mse(const float* sane, const float* test, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
