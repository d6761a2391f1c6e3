/* fwi_serialcalls.h -- C-ABI drop-in for the reference's host launch layer.
 *
 * Each function replaces the same-named launcher of the reference's include/serialcalls.h (C branch,
 * lines 34-55; C++ branch lines 9-31), with the same parameter names, order and meaning, and the same
 * void return.  Arrays are device pointers (float*), x fastest (idx = ix + nx*iz); PF_h is a pinned host
 * pointer to one 2*nx*nz snapshot slot.  The reference's cudaStream_t parameters are opaque stream
 * handles here (fwi_stream_t, pointer-sized: pass a hipStream_t or 0 for the null stream).
 *
 * Semantics are those of the reference kernels in src/parallelcalls.cu, including their arithmetic:
 * stencils are summed in double (the reference's double literals) and rounded to float on store.
 * Ordering follows the reference launch layer: the null stream is the legacy blocking stream, so work on
 * stream1/stream2 and on the null stream orders exactly as src/serialcalls.cu:55-90 relies on.
 *
 * Errors: the reference checks nothing.  These functions stay void and record the first HIP error of
 * the calling thread, readable (and cleared) with fwi_last_error().
 */
#ifndef FWI_SERIALCALLS_H
#define FWI_SERIALCALLS_H

#ifdef __cplusplus
extern "C" {
#endif

typedef void* fwi_stream_t;

/* serialcalls.h:10 / parallelcalls.cu:3-16: v_po = vel^2, v_hor = v_po(1+2 eps), v_nmo = v_po(1+2 delta) */
void pre_variables_h(float* vel, float* epsilon, float* delta, float* v_hor, float* v_po, float* v_nmo, int nx,
                     int nz);
/* serialcalls.h:11 / parallelcalls.cu:18-34: C-PML a_x, b_x for the left and right strips */
void CPML_x_h(float* a_x, float* b_x, int CPML, float Vmax, int nx, float dt, float dh, float f, float d0, float L);
/* serialcalls.h:12 / parallelcalls.cu:36-46: C-PML a_z, b_z for the bottom strip */
void CPML_z_h(float* a_z, float* b_z, int CPML, float Vmax, int nz, float dt, float dh, float f, float d0, float L);
/* serialcalls.h:13 / parallelcalls.cu:48-63: psi_px, psi_fz memory variables on the C-PML cells */
void PSI_F_h(float* P, float* F, float* a_x, float* b_x, float* a_z, float* b_z, float* psi_px, float* psi_fz,
             int nx, int nz, float dh, int CPML);
/* serialcalls.h:14 / parallelcalls.cu:65-86: z_px, z_fz memory variables and aten_p, aten_f */
void ZETA_F_h(float* p2, float* r2, float* P, float* F, float* a_x, float* b_x, float* a_z, float* b_z,
              float* psi_px, float* psi_fz, float* z_px, float* z_fz, float* aten_p, float* aten_f, int nx, int nz,
              float dh, float dh2, int CPML);
/* serialcalls.h:15 / parallelcalls.cu:88-117: coupled VTI leapfrog, source injection, receiver row */
void propagator_F_h(float* p2, float* r2, float* PF, float* P_pre, float* F_pre, float* P_past, float* F_past,
                    float* v_hor, float* v_po, float* v_nmo, float* aten_p, float* aten_f, float* source, float* shot,
                    float C, float dh2, int nx, int nz, int sx, int sz, int it, int CPML);
/* serialcalls.h:16 / parallelcalls.cu:133-142: Res = shot2 - shot1 over nx*nt (the 5th argument is nt) */
void res_h(float* Res, float* shot1, float* shot2, int nx, int nz);
/* serialcalls.h:17 / parallelcalls.cu:120-130: p2 = D2x P, r2 = D2z F */
void SecondDerivate_P_h(float* P_pre, float* F_pre, float* p2, float* r2, float dh2, int nx, int nz);
/* serialcalls.h:18-20 / serialcalls.cu:55-66: one forward step with the snapshot copy of P_past into PF_h */
void launchPropagation(fwi_stream_t stream1, fwi_stream_t stream2, float* PF_h, float* PF, float* P_past,
                       float* P_pre, float* F_past, float* F_pre, float* p2, float* r2, float dh2, int nx, int nz,
                       float* a_x, float* b_x, float* a_z, float* b_z, float* psi_px, float* psi_fz, float dh,
                       int CPML, float* z_px, float* z_fz, float* aten_p, float* aten_f, float* v_hor, float* v_po,
                       float* v_nmo, float* source, float* shot2, float C, int sx, int sz, int it);
/* serialcalls.h:21-24 / serialcalls.cu:68-90: one adjoint step with the snapshot copy and the gradient */
void launchBackPropagation(fwi_stream_t stream1, fwi_stream_t stream2, int nx, int nz, int nt, float* PF_h,
                           float* P_pre, float* P_past, float* F_pre, float* F_past, float* aux1, float* aux2,
                           float* aux3, float* aux4, float* v_hor, float* v_po, float* v_nmo, float* p2cx,
                           float* p2cz, float* r2dx, float* r2cz, float dh2, float* P_pre2, float* F_pre2,
                           float* a_x, float* b_x, float* a_z, float* b_z, float* psixcx, float* psixcz,
                           float* psizdx, float* psizcz, float dh, int CPML, float* zitaxcx, float* zitaxcz,
                           float* zitazdx, float* zitazcz, float* atenxcx, float* atenxcz, float* atenzdx,
                           float* atenzcz, float* Res, float C, float* PF, int it, int sx, int sz, float dt,
                           float* gcx, float* gcz, float* gdx);
/* serialcalls.h:25 / parallelcalls.cu:261-271: g /= norm on rows iz > 15 */
void NORMAL_h(float* gcx, float* gcz, float* gdx, int nx, int nz, float ngcx, float ngcz, float ngdx);
/* serialcalls.h:26 / parallelcalls.cu:273-286: steepest-descent step with v_hor >= v_nmo */
void UpdateModel_h(float* gcx, float* gcz, float* gdx, float* v_hor, float* v_po, float* v_nmo, float alpha1,
                   float alpha2, float alpha3, int nx, int nz);
/* serialcalls.h:27 / parallelcalls.cu:288-302: L-BFGS history slot k: s = m - m_back, y = g - g_back */
void compute_sy_h(float* s_vp, float* s_vn, float* s_vh, float* y_b1, float* y_b2, float* y_b3, float* vp, float* vn,
                  float* vh, float* back_vp, float* back_vn, float* back_vh, float* b1, float* b2, float* b3,
                  float* back_b1, float* back_b2, float* back_b3, int nx, int nz, int k);
/* serialcalls.h:28 / parallelcalls.cu:304-315: q -= alpha_l y_l */
void compute_qi_h(float* q_b1, float* q_b2, float* q_b3, float* y_b1, float* y_b2, float* y_b3, float epsilon_v,
                  float epsilon_d, float epsilon_e, int nx, int nz, int k);
/* serialcalls.h:29 / parallelcalls.cu:317-328: r = gamma q */
void compute_ri_h(float* r_b1, float* r_b2, float* r_b3, float* q_b1, float* q_b2, float* q_b3, float gamma_v,
                  float gamma_d, float gamma_e, int nx, int nz);
/* serialcalls.h:30 / parallelcalls.cu:330-341: r += (alpha_l - beta) s_l */
void update_r_h(float* r_b1, float* r_b2, float* r_b3, float* s_vp, float* s_vn, float* s_vh, float epsilon_v,
                float epsilon_d, float epsilon_e, float beta_v, float beta_d, float beta_e, int nx, int nz, int k);
/* serialcalls.h:31 / parallelcalls.cu:343-404: trial model m = m_back - alpha r with optional clamps */
void update_LBFGS_h(float* r_b1, float* r_b2, float* r_b3, float* vp, float* vn, float* vh, float* back_vp,
                    float* back_vn, float* back_vh, int nx, int nz, float alpha_v, float alpha_d, float alpha_e,
                    int flag_v, int flag_d, int flag_e, int flag_min, int flag_max);

/* No counterpart in the reference (which checks no errors): the first HIP error recorded on this thread by
 * a void launcher above, as a hipError_t value (0 = none); reading it clears it. */
int fwi_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* FWI_SERIALCALLS_H */
