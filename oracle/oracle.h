/*
 * CPU oracle for the NeuronGPU per-timestep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used by tests/ as the checker, by __graft_entry__.smoke() as the
 * checker, and by bench.py's cpu_baseline leg as the timed CPU baseline.  The
 * product library (neurongpu_amd/libneurongpu.so) never links, loads or calls
 * anything in this directory.
 *
 * One source, two builds: liboracle_f32.so (REAL = float; arithmetic op
 * sequence identical to the HIP kernels, so results can be compared bit for
 * bit) and liboracle_f64.so (REAL = double; the accuracy reference).
 *
 * What it restates (reference file:line):
 *   Cash-Karp RK5 step + controller      src/rk5.h:57-162, src/rk5.cu:23-60
 *   aeif_cond_beta RHS / spike / reset   src/aeif_cond_beta_kernel.h:153-214
 *   aeif_cond_beta defaults, g0 norm.    src/aeif_cond_beta.cu:25-87, 153-160
 *   aeif_cond_alpha RHS / g0             src/aeif_cond_alpha_kernel.h:149-211,
 *                                        src/aeif_cond_alpha.cu:57-69
 *   iaf_psc_exp propagators / update     src/iaf_psc_exp.cu:53-119
 *   step phase order                     src/neurongpu.cu:448-605
 *   spike buffer / delay semantics       src/spike_buffer.cu:143-302,
 *                                        src/connect.cu:26-62
 *   delivery + pre-synaptic STDP         src/get_spike.cu:48-85, 116-140
 *   post-synaptic STDP                   src/rev_spike.cu:41-147
 *   STDP weight rule                     src/stdp.cu:24-51
 *   Poisson direct connections           src/poiss_gen.cu:43-74,
 *                                        src/neurongpu.cu:1156-1192
 *   spike generator                      src/spike_generator.cu:37-55, 209-238
 *   iaf_psc_exp_g / _hc                  src/iaf_psc_exp_g.cu:34-61, 136-150,
 *                                        src/iaf_psc_exp_hc.cu:29-53
 *   izhikevich (Euler), _psc_exp(_2s,_5s) src/izhikevich.cu:35-72,
 *                                        src/izhikevich_psc_exp_2s.cu:41-83
 *   ext_neuron                           src/ext_neuron.cu:23-39
 *   CreatePoissonGenerator nodes         src/poisson.cu:27-54, 98-147,
 *                                        src/neurongpu.cu:245-265, 461-470
 *
 * Deliberate restatement choices (shared with the HIP kernels):
 *   - reciprocals of Delta_T, C_m, tau_w, tau_rise, tau_decay are computed
 *     once at calibration and multiplied (the reference divides);
 *   - exp/log in FP32 are this project's own Cody-Waite + polynomial
 *     functions (orc_expf/orc_logf), so GPU and CPU agree bit for bit;
 *   - Poisson counts come from Threefry2x32-20 (Salmon et al., SC'11) and
 *     inversion by sequential search over a per-rate CDF table of 32-bit
 *     thresholds, instead of curand (closed source);
 *   - the step-size controller's err^-0.2 and err^-0.25 come from one
 *     polynomial sequence (orc_pow_ctrl) instead of powf.
 */
#ifndef NGPU_ORACLE_H
#define NGPU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_AEIF_COND_BETA = 1,
  ORC_AEIF_COND_ALPHA = 2,
  ORC_IAF_PSC_EXP = 3,
  ORC_SPIKE_GENERATOR = 4,
  ORC_POISSON_GENERATOR = 5,
  ORC_PARROT_NEURON = 6,
  ORC_AEIF_PSC_ALPHA = 7,
  ORC_AEIF_PSC_EXP = 8,
  ORC_AEIF_PSC_DELTA = 9,
  ORC_IZH_COND_BETA = 10,
  ORC_POISSON_LEGACY = 11,
  ORC_IAF_PSC_EXP_G = 12,
  ORC_IAF_PSC_EXP_HC = 13,
  ORC_IZHIKEVICH = 14,
  ORC_IZH_PSC_EXP = 15,
  ORC_IZH_PSC_EXP_2S = 16,
  ORC_IZH_PSC_EXP_5S = 17,
  ORC_EXT_NEURON = 18
};

/* ---- primitives ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
int orc_poisson_table(double lambda, uint32_t *tab, int max_len);
int orc_poisson_sample(const uint32_t *tab, uint32_t u);
float orc_expf(float x);
float orc_logf(float x);
float orc_pow_ctrl(float e, float c);
float orc_stdp_update(float w, float Dt, const float *param7);
int orc_real_size(void);

/* single-neuron calibrations */
double orc_beta_g0(double tau_rise, double tau_decay);
void orc_iaf_propagators(double h, double tau_m, double C_m, double tau_ex,
                         double tau_in, float out6[6]);

/* fixed-step and adaptive RK5 on y' = lambda*y (order / tolerance checks) */
double orc_rk5_linear_fixed(double lambda, double y0, double T, int n_steps);
double orc_rk5_linear_adaptive(double lambda, double y0, double T, double h0,
                               double h_min, double eps, long *n_acc,
                               long *n_rej);

/* ---- network ---- */
typedef struct orc_net orc_net;

orc_net *orc_net_create(double dt, double t0, uint64_t seed);
void orc_net_free(orc_net *net);
int orc_net_add_group(orc_net *net, int model, int n, int n_port);
int orc_net_n_var(orc_net *net, int group);
int orc_net_n_param(orc_net *net, int group);
/* SoA [idx][n] arrays of the group in REAL precision */
void orc_net_set_var(orc_net *net, int group, int idx, const void *vals);
void orc_net_get_var(orc_net *net, int group, int idx, void *vals);
void orc_net_set_param(orc_net *net, int group, int idx, const void *vals);
void orc_net_get_param(orc_net *net, int group, int idx, void *vals);
/* h_min_rel, h0_rel, rk5 tolerance */
void orc_net_set_rk5(orc_net *net, int group, double h_min_rel, double h0_rel,
                     double eps);
void orc_net_spike_gen(orc_net *net, int node, int n, const float *times_ms,
                       const float *heights);
void orc_net_poisson(orc_net *net, int node, float rate, float origin,
                     float start, float stop);
/* rate (Hz) of an ORC_POISSON_LEGACY group (CreatePoissonGenerator);
 * returns -1 if the mean count per step is negative or above the table */
int orc_net_legacy_rate(orc_net *net, int group, float rate);
int orc_net_syn_group_stdp(orc_net *net, const float *param7);
void orc_net_connect(orc_net *net, int n, const int *src, const int *tgt,
                     const uint8_t *port, const uint8_t *syn_group,
                     const float *weight, const float *delay);
int orc_net_record(orc_net *net, int node, int var_idx, int is_param);
int orc_net_calibrate(orc_net *net);
int orc_net_run(orc_net *net, long n_steps, int n_threads);

long orc_net_n_spikes(orc_net *net);
/* spike list: step index (0-based since calibration), node, multiplicity */
void orc_net_get_spikes(orc_net *net, long *step, int *node, int *count);
/* summed height of each entry of the spike list */
void orc_net_get_spike_heights(orc_net *net, double *height);
int orc_net_n_records(orc_net *net);
long orc_net_n_record_rows(orc_net *net);
void orc_net_get_records(orc_net *net, double *out); /* rows x (1+n_rec) */
long orc_net_n_conn(orc_net *net);
/* canonical order: source, delay, insertion */
void orc_net_get_conn(orc_net *net, int *src, int *tgt, int *port, int *delay,
                      float *weight);
void orc_net_counters(orc_net *net, long *out4); /* acc, rej, events, poisson */

#ifdef __cplusplus
}
#endif
#endif
