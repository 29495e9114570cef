// Node model descriptors: variable/parameter names in the order of the
// reference enums, and NodeInit defaults.
//   aeif_cond_beta   src/aeif_cond_beta_kernel.h:30-112, aeif_cond_beta.cu:25-56
//   aeif_cond_alpha  src/aeif_cond_alpha_kernel.h:30-112, aeif_cond_alpha.cu:25-55
//   iaf_psc_exp      src/iaf_psc_exp.h:32-83, iaf_psc_exp.cu:134-170
//   spike_generator  src/spike_generator.cu:26-35
//   poisson_generator src/poiss_gen_variables.h, src/poiss_gen.cu:77-95
//   spike_detector   src/spike_detector.cu:26-38, 72-80
//   parrot_neuron    src/parrot_neuron.cu:26-46, 75-100
//   aeif_psc_alpha   src/aeif_psc_alpha_kernel.h:30-105, aeif_psc_alpha.cu:24-48
//   aeif_psc_exp     src/aeif_psc_exp_kernel.h:30-100, aeif_psc_exp.cu:24-46
//   aeif_psc_delta   src/aeif_psc_delta_kernel.h:30-95, aeif_psc_delta.cu:24-45
//   izhikevich_cond_beta  src/izhikevich_cond_beta_kernel.h:30-98,
//                    izhikevich_cond_beta.cu:24-50
//   user_m1, user_m2 src/user_m1_kernel.h, src/user_m2_kernel.h (the
//                    aeif_cond_beta model under other names)
// The group parameter "rk5_tol" (RK5 error tolerance) has no reference
// counterpart: the reference hard-codes 1e-6 (src/rk5.cu:55); its default
// here is the same 1e-6.
#include <algorithm>

#include "engine.h"

namespace ngpu {

namespace {

std::vector<ModelDesc> MakeModels() {
  std::vector<ModelDesc> m(kNumModels);
  ModelDesc &b = m[kAeifCondBeta];
  b.name = "aeif_cond_beta";
  b.scal_var = {"V_m", "w"};
  b.port_var = {"g", "g1"};
  b.scal_param = {"V_th",    "Delta_T", "g_L",    "E_L",   "C_m",
                  "a",       "b",       "tau_w",  "I_e",   "V_peak",
                  "V_reset", "t_ref",   "refractory_step", "den_delay"};
  b.port_param = {"E_rev", "tau_rise", "tau_decay", "g0"};
  b.group_param = {"h_min_rel", "h0_rel", "rk5_tol"};
  b.scal_param_default = {-50.4f, 2.0f, 30.0f, -70.6f, 281.0f, 4.0f, 80.5f,
                          144.0f, 0.0f,  0.0f, -60.0f, 0.0f,   0.0f, 0.0f};
  b.port_param_default = {0.0f, 2.0f, 20.0f, 0.0f};
  b.group_param_default = {1.0e-3f, 1.0e-2f, 1.0e-6f};
  b.scal_var_default = {-70.6f, 0.0f};
  b.port_var_default = {0.0f, 0.0f};
  b.fixed_ports = -1;

  ModelDesc &a = m[kAeifCondAlpha];
  a = b;
  a.name = "aeif_cond_alpha";
  a.port_param = {"E_rev", "tau_syn", "g0"};
  a.port_param_default = {0.0f, 2.0f, 0.0f};

  ModelDesc &f = m[kIafPscExp];
  f.name = "iaf_psc_exp";
  f.scal_var = {"I_syn_ex", "I_syn_in", "V_m_rel", "refractory_step"};
  f.scal_param = {"tau_m",   "C_m",    "E_L",   "I_e",   "Theta_rel",
                  "V_reset_rel", "tau_ex", "tau_in", "t_ref", "den_delay",
                  "P20",     "P11ex",  "P11in", "P21ex", "P21in", "P22"};
  f.scal_param_default = {10.0f, 250.0f, -70.0f, 0.0f, 15.0f, 0.0f,
                          2.0f,  2.0f,   2.0f,   0.0f, 0.0f,  0.0f,
                          0.0f,  0.0f,   0.0f,   0.0f};
  f.scal_var_default = {0.0f, 0.0f, 0.0f, 0.0f};
  f.fixed_ports = 2;

  ModelDesc &s = m[kSpikeGenerator];
  s.name = "spike_generator";
  s.array_param = {"spike_times", "spike_heights"};
  s.fixed_ports = 0;

  ModelDesc &p = m[kPoissonGenerator];
  p.name = "poisson_generator";
  p.scal_param = {"rate", "origin", "start", "stop"};
  p.scal_param_default = {0.0f, 0.0f, 0.0f, 1.0e30f};
  p.fixed_ports = 0;

  ModelDesc &d = m[kSpikeDetector];
  d.name = "spike_detector";
  d.scal_var = {"input_spike_height", "spike_height"};
  d.scal_param = {"hold_spike_height"};
  d.scal_param_default = {1.0f};
  d.scal_var_default = {0.0f, 0.0f};
  d.fixed_ports = 1;

  ModelDesc &r = m[kParrotNeuron];
  r.name = "parrot_neuron";
  r.scal_var = {"input_spike_height", "dummy_input", "V"};
  r.scal_param = {"hold_spike_height", "den_delay"};
  r.scal_param_default = {0.0f, 0.0f};
  r.scal_var_default = {0.0f, 0.0f, 0.0f};
  r.fixed_ports = 2;

  ModelDesc &pa = m[kAeifPscAlpha];
  pa = b;
  pa.name = "aeif_psc_alpha";
  pa.port_var = {"I_syn", "I1_syn"};
  pa.port_param = {"tau_syn", "I0"};
  pa.port_param_default = {0.2f, 0.0f};

  ModelDesc &pe = m[kAeifPscExp];
  pe = b;
  pe.name = "aeif_psc_exp";
  pe.port_var = {"I_syn"};
  pe.port_param = {"tau_syn"};
  pe.port_param_default = {0.2f};
  pe.port_var_default = {0.0f};

  ModelDesc &pd = m[kAeifPscDelta];
  pd = b;
  pd.name = "aeif_psc_delta";
  pd.port_var = {};
  pd.port_param = {};
  pd.port_param_default = {};
  pd.port_var_default = {};

  ModelDesc &iz = m[kIzhikevichCondBeta];
  iz.name = "izhikevich_cond_beta";
  iz.scal_var = {"V_m", "u"};
  iz.port_var = {"g", "g1"};
  iz.scal_param = {"V_th", "a", "b", "c", "d", "I_e", "t_ref",
                   "refractory_step", "den_delay"};
  iz.port_param = {"E_rev", "tau_rise", "tau_decay", "g0"};
  iz.group_param = {"h_min_rel", "h0_rel", "rk5_tol"};
  iz.scal_param_default = {30.0f, 0.02f, 0.2f, -65.0f, 8.0f, 0.0f, 0.0f,
                           0.0f, 0.0f};
  iz.port_param_default = {0.0f, 2.0f, 20.0f, 0.0f};
  iz.group_param_default = {1.0e-3f, 1.0e-2f, 1.0e-6f};
  iz.scal_var_default = {-70.0f, 0.2f * -70.0f};  // u = b * V_m
  iz.port_var_default = {0.0f, 0.0f};
  iz.fixed_ports = -1;

  m[kUserM1] = b;
  m[kUserM1].name = "user_m1";
  m[kUserM2] = b;
  m[kUserM2].name = "user_m2";

  // iaf_psc_exp_g (src/iaf_psc_exp_g.h:32-76, iaf_psc_exp_g.cu:96-118)
  ModelDesc &ig = m[kIafPscExpG];
  ig.name = "iaf_psc_exp_g";
  ig.scal_var = {"I_syn", "V_m_rel", "refractory_step"};
  ig.scal_var_default = {0.0f, 0.0f, 0.0f};
  ig.scal_param = {"I_e"};
  ig.scal_param_default = {0.0f};
  ig.group_param = {"tau_m", "C_m", "E_L", "Theta_rel", "V_reset_rel",
                    "tau_syn", "t_ref"};
  ig.group_param_default = {10.0f, 250.0f, -65.0f, 15.0f, 0.0f, 0.5f, 2.0f};
  ig.fixed_ports = 1;

  // iaf_psc_exp_hc (src/iaf_psc_exp_hc.h:42-52): propagators fixed
  ModelDesc &ih = m[kIafPscExpHc];
  ih.name = "iaf_psc_exp_hc";
  ih.scal_var = {"I_syn", "V_m_rel", "refractory_step"};
  ih.scal_var_default = {0.0f, 0.0f, 0.0f};
  ih.scal_param = {"I_e"};
  ih.scal_param_default = {0.0f};
  ih.fixed_ports = 1;

  // izhikevich (src/izhikevich.h:54-73, izhikevich.cu:96-110)
  ModelDesc &ie = m[kIzhikevich];
  ie.name = "izhikevich";
  ie.scal_var = {"I_syn", "V_m", "u", "refractory_step"};
  ie.scal_var_default = {0.0f, -70.0f, -70.0f * 0.2f, 0.0f};
  ie.scal_param = {"I_e", "den_delay"};
  ie.scal_param_default = {0.0f, 0.0f};
  ie.group_param = {"V_th", "a", "b", "c", "d", "t_ref"};
  ie.group_param_default = {30.0f, 0.02f, 0.2f, -65.0f, 8.0f, 0.0f};
  ie.fixed_ports = 1;

  // izhikevich_psc_exp, _2s, _5s (src/izhikevich_psc_exp.h:56-76,
  // izhikevich_psc_exp.cu:100-116)
  const char *izp_name[3] = {"izhikevich_psc_exp", "izhikevich_psc_exp_2s",
                             "izhikevich_psc_exp_5s"};
  const ModelId izp_id[3] = {kIzhikevichPscExp, kIzhikevichPscExp2s,
                             kIzhikevichPscExp5s};
  for (int k = 0; k < 3; k++) {
    ModelDesc &ip = m[izp_id[k]];
    ip = ie;
    ip.name = izp_name[k];
    ip.group_param = {"V_th", "a", "b", "c", "d", "tau_syn", "t_ref"};
    ip.group_param_default = {30.0f, 0.02f, 0.2f, -65.0f, 8.0f, 2.0f, 0.0f};
  }

  // ext_neuron (src/ext_neuron.h:46-60, ext_neuron.cu:45-84)
  ModelDesc &ex = m[kExtNeuron];
  ex.name = "ext_neuron";
  ex.port_var = {"port_input", "port_value"};
  ex.port_var_default = {0.0f, 0.0f};
  ex.scal_param = {"den_delay"};
  ex.scal_param_default = {0.0f};
  ex.port_param = {"port_weight"};
  ex.port_param_default = {1.0f};
  ex.fixed_ports = -1;

  // CreatePoissonGenerator nodes (src/neurongpu.cu:245-265): a BaseNeuron
  // group with no variables, parameters or receptor ports
  ModelDesc &lp = m[kPoissonLegacy];
  lp.name = "base_neuron";
  lp.fixed_ports = 0;
  return m;
}

}  // namespace

const ModelDesc &Model(ModelId id) {
  static const std::vector<ModelDesc> models = MakeModels();
  return models[id];
}

int FindModel(const std::string &name) {
  for (int i = 0; i < kNumModels; i++)
    if (i != kPoissonLegacy && name == Model((ModelId)i).name) return i;
  return -1;
}

int Group::VarIdx(const std::string &name, int port) const {
  const ModelDesc &m = Model(model);
  auto it = std::find(m.scal_var.begin(), m.scal_var.end(), name);
  if (it != m.scal_var.end()) return (int)(it - m.scal_var.begin());
  auto jt = std::find(m.port_var.begin(), m.port_var.end(), name);
  if (jt != m.port_var.end()) {
    if (port < 0 || port >= n_port) throw ngpu_exception("Port index out of range");
    return (int)m.scal_var.size() + (int)m.port_var.size() * port +
           (int)(jt - m.port_var.begin());
  }
  return -1;
}

int Group::ParamIdx(const std::string &name, int port) const {
  const ModelDesc &m = Model(model);
  auto it = std::find(m.scal_param.begin(), m.scal_param.end(), name);
  if (it != m.scal_param.end()) return (int)(it - m.scal_param.begin());
  auto jt = std::find(m.port_param.begin(), m.port_param.end(), name);
  if (jt != m.port_param.end()) {
    if (port < 0 || port >= n_port) throw ngpu_exception("Port index out of range");
    return (int)m.scal_param.size() + (int)m.port_param.size() * port +
           (int)(jt - m.port_param.begin());
  }
  return -1;
}

Group::~Group() {
  for (auto *b : d_derived) delete b;
}

}  // namespace ngpu
