// app.hpp -- the settings-file driven application (JSON or TOML).
//
// Drop-in for pfc::ui::App<ConcreteModel> (include/openpfc/frontend/ui/
// app.hpp:52-170) on the HIP stack: `App<M> app(argc, argv); app.main();`
// reads the settings file named by argv[1] (.json or .toml), builds the World
// ("domain", from_json_world_time.hpp:42-184), the Time ("timestepping",
// :190-284), the FFT ("plan_options.backend": "hip" or "rocm"; the default),
// the model (constructed from the FFT and the World, "model.params" applied
// as the reference's from_json does), the initial and boundary conditions
// ("initial_conditions" / "boundary_conditions": constant, single_seed,
// seed_grid, random_seeds, from_file, uniform_noise, fixed, moving;
// field_modifier_registry.hpp:108-118), the
// results writers ("fields": [{name, data}] written with BinaryWriter at
// "saveat", simulation_wiring_writers.hpp) and the "simulator" section
// (result_counter, increment: simulation_wiring_simulator_section.hpp), then
// runs the Simulator loop.  Device-resident models stay on the device for
// the whole run: initial and boundary conditions and writers use the
// model's device field.  A run restarted at increment > 0 applies its
// initial conditions once before the loop, so a from_file condition loads
// the checkpoint.
#pragma once

#include "io.hpp"
#include "json.hpp"
#include "pfc_model_hip.hpp"
#include "simulator.hpp"
#include "time.hpp"
#include "toml.hpp"
#include "tungsten_model_hip.hpp"

#include <chrono>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>

namespace pfc::ui {

using json::Value;

// settings_loader.hpp: by file extension
inline Value load_settings_file(const std::string &path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open settings file '" + path + "'");
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string ext = std::filesystem::path(path).extension().string();
  if (ext == ".json") return json::parse(ss.str());
  if (ext == ".toml") return toml::parse(ss.str());
  throw std::invalid_argument("settings file '" + path + "': extension must be .json or .toml");
}

// json_helpers.hpp get_json_value: a key at the top level or in `section`
inline const Value *get_value(const Value &j, const std::string &key, const std::string &section) {
  if (const Value *v = j.find(key)) return v;
  if (const Value *s = j.find(section))
    if (const Value *v = s->find(key)) return v;
  return nullptr;
}

inline std::string config_error(const std::string &key, const std::string &what,
                                const std::string &expected, const Value *got,
                                const std::string &example) {
  return "Invalid configuration: '" + key + "' (" + what + ") must be a " + expected +
         ", got " + (got ? got->dump() : std::string("missing")) + ". Example: " + example;
}

// from_json<World> (from_json_world_time.hpp:42-184)
inline World world_from_json(const Value &j) {
  int n[3];
  double d[3];
  const char *nk[3] = {"Lx", "Ly", "Lz"}, *dk[3] = {"dx", "dy", "dz"}, *ax = "XYZ";
  for (int i = 0; i < 3; ++i) {
    const Value *v = get_value(j, nk[i], "domain");
    if (!v || !v->is_number_integer() || v->as_int() <= 0)
      throw std::invalid_argument(config_error(nk[i], std::string("number of grid points in ") +
                                                          ax[i] + " direction",
                                               "positive integer", v,
                                               std::string("\"") + nk[i] + "\": 256"));
    n[i] = (int)v->as_int();
    const Value *s = get_value(j, dk[i], "domain");
    if (!s || !s->is_number() || !std::isfinite(s->as_double()) || s->as_double() <= 0)
      throw std::invalid_argument(config_error(dk[i], std::string("grid spacing in ") + ax[i] +
                                                          " direction",
                                               "positive float", s,
                                               std::string("\"") + dk[i] + "\": 1.0"));
    d[i] = s->as_double();
  }
  const Value *o = get_value(j, "origin", "domain");
  if (!o) o = get_value(j, "origo", "domain");
  if (!o || !o->is_string() || (o->as_string() != "center" && o->as_string() != "corner"))
    throw std::invalid_argument(config_error("origin", "coordinate system origin",
                                             "string ('center' or 'corner')", o,
                                             "\"origin\": \"center\""));
  Real3 x0{0, 0, 0};
  if (o->as_string() == "center")
    for (int i = 0; i < 3; ++i) x0[i] = -0.5 * d[i] * n[i];
  return domain::create(GridSize({n[0], n[1], n[2]}), PhysicalOrigin(x0),
                        GridSpacing({d[0], d[1], d[2]}));
}

// from_json<Time> (from_json_world_time.hpp:190-284)
inline Time time_from_json(const Value &j) {
  const char *keys[4] = {"t0", "t1", "dt", "saveat"};
  double v[4];
  for (int i = 0; i < 4; ++i) {
    const Value *x = get_value(j, keys[i], "timestepping");
    if (!x)
      throw std::invalid_argument("Missing required time stepping parameters (t0, t1, dt, saveat)");
    if (!x->is_number() || !std::isfinite(x->as_double()))
      throw std::invalid_argument(config_error(keys[i], "time stepping parameter", "finite float",
                                               x, std::string("\"") + keys[i] + "\": 1.0"));
    v[i] = x->as_double();
  }
  if (v[1] <= v[0])
    throw std::invalid_argument(config_error("t1", "simulation end time", "value greater than t0",
                                             get_value(j, "t1", "timestepping"), "\"t1\": 100.0"));
  if (v[2] <= 0)
    throw std::invalid_argument(config_error("dt", "timestep size", "positive float",
                                             get_value(j, "dt", "timestepping"), "\"dt\": 0.01"));
  return Time({v[0], v[1], v[2]}, v[3]);
}

// plan_options.backend (backend_from_string.hpp:24-40,
// from_json_fft_backend.hpp:38-75).  The models of this App are
// device-resident and transform on the device whatever the base FFT is, as
// the reference's tungsten_hip does (its TungstenHIP builds a rocFFT plan of
// its own next to the FFTW base FFT of tungsten_performance.json), so "fftw"
// is accepted beside "hip"/"rocm"; "cuda" is not built here and unknown names
// are refused.
inline void check_backend(const Value &j) {
  const Value *po = j.find("plan_options");
  if (!po) return;
  const Value *b = po->find("backend");
  if (!b) return;
  if (!b->is_string())
    throw std::invalid_argument("plan_options.backend must be a string, got " + b->dump());
  const fft::Backend be = fft::backend_from_string(b->as_string());
  if (be == fft::Backend::CUDA)
    throw std::runtime_error("CUDA FFT backend requested, but this library provides the HIP "
                             "stack only");
}

inline double num(const Value &j, const std::string &key, const std::string &ctx) {
  const Value *v = j.find(key);
  if (!v || !v->is_number())
    throw std::invalid_argument(ctx + ": missing or non-numeric '" + key + "'");
  return v->as_double();
}

// A required numeric field of a modifier, with the reference's message
// ("Invalid JSON input: missing or invalid 'X0' field.").
inline const Value &required_number(const Value &j, const std::string &key) {
  const Value *v = j.find(key);
  if (!v || !v->is_number())
    throw std::invalid_argument("Invalid JSON input: missing or invalid '" + key + "' field.");
  return *v;
}

// Field modifiers by "type" (field_modifier_registry.hpp:108-118,
// from_json_field_modifiers.hpp).
inline std::unique_ptr<FieldModifier> field_modifier_from_json(const Value &j) {
  const Value *t = j.find("type");
  if (!t || !t->is_string())
    throw std::invalid_argument("field modifier: missing string 'type' in " + j.dump());
  const std::string type = t->as_string();
  std::unique_ptr<FieldModifier> m;
  if (type == "constant") {
    m = std::make_unique<Constant>(num(j, "n0", "constant"));
  } else if (type == "single_seed") {
    auto s = std::make_unique<SingleSeed>();
    s->set_amplitude(num(j, "amp_eq", "single_seed"));
    s->set_density(num(j, "rho_seed", "single_seed"));
    if (j.contains("radius")) s->set_radius(num(j, "radius", "single_seed"));
    m = std::move(s);
  } else if (type == "from_file") {
    const Value *f = j.find("filename");
    if (!f || !f->is_string())
      throw std::invalid_argument("from_file: missing string 'filename'");
    m = std::make_unique<FileReader>(f->as_string());
  } else if (type == "uniform_noise") {
    const Value *seed = j.find("seed");
    m = std::make_unique<UniformNoise>(seed && seed->is_number_integer() ? (std::uint64_t)seed->as_int()
                                                                          : 0,
                                       num(j, "low", "uniform_noise"),
                                       num(j, "high", "uniform_noise"));
  } else if (type == "fixed") {
    m = std::make_unique<FixedBC>(num(j, "rho_low", "fixed"), num(j, "rho_high", "fixed"));
  } else if (type == "seed_grid") {
    // from_json_field_modifiers.hpp:85-125
    auto g = std::make_unique<SeedGrid>();
    g->set_Ny((int)required_number(j, "Ny").as_double());
    g->set_Nz((int)required_number(j, "Nz").as_double());
    g->set_X0(required_number(j, "X0").as_double());
    g->set_radius(required_number(j, "radius").as_double());
    g->set_amplitude(required_number(j, "amplitude").as_double());
    g->set_density(required_number(j, "rho").as_double());
    m = std::move(g);
  } else if (type == "random_seeds") {
    // from_json_field_modifiers.hpp:65-83
    auto r = std::make_unique<RandomSeeds>();
    r->set_amplitude(required_number(j, "amplitude").as_double());
    r->set_density(required_number(j, "rho").as_double());
    m = std::move(r);
  } else if (type == "moving") {
    // from_json_field_modifiers.hpp:157-197 (threshold keeps its default 0.1)
    auto b = std::make_unique<MovingBC>();
    b->set_rho_low(required_number(j, "rho_low").as_double());
    b->set_rho_high(required_number(j, "rho_high").as_double());
    b->set_xwidth(required_number(j, "width").as_double());
    b->set_alpha(required_number(j, "alpha").as_double());
    b->set_disp(required_number(j, "disp").as_double());
    b->set_xpos(required_number(j, "xpos").as_double());
    m = std::move(b);
  } else {
    throw std::invalid_argument("unknown field modifier type '" + type +
                                "' (known: constant, single_seed, seed_grid, random_seeds, "
                                "from_file, uniform_noise, fixed, moving)");
  }
  if (const Value *tg = j.find("target")) m->set_field_name(tg->as_string());
  return m;
}

// Model construction and parameters per model type.
template <typename M> struct ModelTraits;

template <typename T> struct ModelTraits<TungstenHIP<T>> {
  static constexpr int precision = std::is_same_v<T, double> ? PFC_HIP_F64 : PFC_HIP_F32;
  static const char *name() { return "tungsten"; }
  // tungsten_input.hpp:399-440 (every base parameter required)
  static std::unique_ptr<TungstenHIP<T>> create(fft::HipFft &f, const World &w, const Value &p) {
    auto m = std::make_unique<TungstenHIP<T>>(f, w);
    TungstenParams &q = m->params;
    const char *ctx = "model.params (tungsten)";
    q.n0 = num(p, "n0", ctx);
    q.n_sol = num(p, "n_sol", ctx);
    q.n_vap = num(p, "n_vap", ctx);
    q.T = num(p, "T", ctx);
    q.T0 = num(p, "T0", ctx);
    q.Bx = num(p, "Bx", ctx);
    q.alpha = num(p, "alpha", ctx);
    q.alpha_farTol = num(p, "alpha_farTol", ctx);
    const Value *ho = p.find("alpha_highOrd");
    if (!ho || !ho->is_number_integer())
      throw std::invalid_argument(std::string(ctx) + ": 'alpha_highOrd' must be an integer");
    q.alpha_highOrd = (int)ho->as_int();
    q.lambda = num(p, "lambda", ctx);
    q.stabP = num(p, "stabP", ctx);
    q.shift_u = num(p, "shift_u", ctx);
    q.shift_s = num(p, "shift_s", ctx);
    q.p2 = num(p, "p2", ctx);
    q.p3 = num(p, "p3", ctx);
    q.p4 = num(p, "p4", ctx);
    q.q20 = num(p, "q20", ctx);
    q.q21 = num(p, "q21", ctx);
    q.q30 = num(p, "q30", ctx);
    q.q31 = num(p, "q31", ctx);
    q.q40 = num(p, "q40", ctx);
    return m;
  }
};

template <> struct ModelTraits<PFCModelHIP> {
  static constexpr int precision = PFC_HIP_F64;
  static const char *name() { return "pfc"; }
  static std::unique_ptr<PFCModelHIP> create(fft::HipFft &f, const World &w, const Value &p) {
    return std::make_unique<PFCModelHIP>(f, w, num(p, "epsilon", "model.params (pfc)"));
  }
};

template <class ConcreteModel> class App {
public:
  App(int argc, char **argv) {
    if (argc <= 1)
      throw std::runtime_error("OpenPFC App: configuration file required (pass path to JSON "
                               "or TOML as first argument)");
    try {
      m_settings = load_settings_file(argv[1]);
    } catch (const std::exception &e) {
      throw std::runtime_error(std::string("OpenPFC App: failed to load settings: ") + e.what());
    }
  }
  explicit App(Value settings) : m_settings(std::move(settings)) {}

  const Value &settings() const { return m_settings; }

  int main() {
    using clock = std::chrono::steady_clock;
    check_backend(m_settings);
    World world = world_from_json(m_settings);
    Time time = time_from_json(m_settings);
    fft::HipFft fft(world, ModelTraits<ConcreteModel>::precision);
    Value params = Value::object();
    if (const Value *mj = m_settings.find("model"))
      if (const Value *p = mj->find("params")) params = *p;
    auto model = ModelTraits<ConcreteModel>::create(fft, world, params);
    const auto t_init = clock::now();
    model->initialize(time.get_dt());
    Simulator sim(*model, time);
    // simulator section: result_counter + 1, increment
    if (const Value *s = m_settings.find("simulator")) {
      if (const Value *rc = s->find("result_counter")) {
        if (!rc->is_number_integer())
          throw std::invalid_argument("Invalid JSON input: missing or invalid 'result_counter' field.");
        sim.set_result_counter((int)rc->as_int() + 1);
      }
      if (const Value *inc = s->find("increment")) {
        if (!inc->is_number_integer())
          throw std::invalid_argument("Invalid JSON input: missing or invalid 'increment' field.");
        time.set_increment((int)inc->as_int());
      }
    }
    if (const Value *ics = m_settings.find("initial_conditions"))
      for (const Value &j : ics->items()) sim.add_initial_conditions(field_modifier_from_json(j));
    if (const Value *bcs = m_settings.find("boundary_conditions"))
      for (const Value &j : bcs->items()) sim.add_boundary_conditions(field_modifier_from_json(j));
    add_writers(sim, world, fft);
    if (time.get_increment() > 0) sim.apply_initial_conditions();
    std::printf("[app] %s model, %dx%dx%d, t %.6g -> %.6g, dt %.6g, setup %.3f s\n",
                ModelTraits<ConcreteModel>::name(), world.size[0], world.size[1], world.size[2],
                time.get_current(), time.get_t1(), time.get_dt(),
                std::chrono::duration<double>(clock::now() - t_init).count());
    std::fflush(stdout);
    const auto t0 = clock::now();
    int steps = 0;
    // steps between save points in one model call (fused on the device)
    while (!done(sim)) steps += sim.step_block(1 << 30);
    pfc::detail::check_hip(hipStreamSynchronize(fft.stream()), "App: final sync");
    const double wall = std::chrono::duration<double>(clock::now() - t0).count();
    std::printf("[app] %d steps in %.3f s (%.3f ms/step), t = %.6g, %d result sets\n", steps,
                wall, steps ? 1e3 * wall / steps : 0.0, time.get_current(),
                sim.get_result_counter());
    std::fflush(stdout);
    return 0;
  }

private:
  void add_writers(Simulator &sim, const World &world, fft::HipFft &fft) {
    const Value *fields = m_settings.find("fields");
    const Value *saveat = get_value(m_settings, "saveat", "timestepping");
    if (!fields || !saveat || !(saveat->as_double() > 0)) return;
    for (const Value &f : fields->items()) {
      const std::string name = f.at("name").as_string();
      const std::string data = f.at("data").as_string();
      if (const Value *w = f.find("writer"); w && w->as_string() != "binary")
        throw std::invalid_argument("fields: writer '" + w->as_string() +
                                    "' is not available (binary only)");
      const std::filesystem::path parent = std::filesystem::path(data).parent_path();
      if (!parent.empty()) std::filesystem::create_directories(parent);
      const auto inbox = fft::get_inbox(fft);
      const std::array<int, 3> g{world.size[0], world.size[1], world.size[2]};
      const std::array<int, 3> l{inbox.size[0], inbox.size[1], inbox.size[2]};
      const std::array<int, 3> o{inbox.low[0], inbox.low[1], inbox.low[2]};
      sim.add_results_writer([=](Model &m, const Time &, int file_num) {
        const std::string path = utils::format_with_number(data, file_num);
        DeviceField d;
        if (m.get_device_field(name, d)) {
          pfc::detail::check_abi(pfc_hip_field_write(path.c_str(), d.ptr, d.precision,
                                                     PFC_HIP_IO_REAL, g.data(), l.data(),
                                                     o.data(), 1, d.stream),
                                 "App results writer");
        } else {
          m.prepare_for_field_modifiers();
          BinaryWriter w(data);
          w.set_domain(g, l, o);
          w.write(file_num, m.get_real_field(name));
        }
      });
    }
  }

  Value m_settings;
};

} // namespace pfc::ui
