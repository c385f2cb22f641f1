// openpfc_amd_app -- settings-file driven runs on the HIP stack.
//
//   openpfc_amd_app config.json|config.toml
//
// model.name selects the model: "tungsten" (TungstenHIP<double>; with
// plan_options.precision "float", TungstenHIP<float>) or "pfc"
// (PFCModelHIP, params {epsilon}).  Mirrors the reference's tungsten
// application entry (apps/tungsten/include/tungsten/common/
// tungsten_app_main.hpp): errors are printed and the exit status is 1.
#include <openpfc_amd/app.hpp>

#include <cstdlib>
#include <iostream>

int main(int argc, char **argv) {
  try {
    if (argc <= 1) {
      std::cerr << "usage: " << argv[0] << " <config.json|config.toml>\n";
      return EXIT_FAILURE;
    }
    const pfc::json::Value s = pfc::ui::load_settings_file(argv[1]);
    std::string name = "tungsten", precision = "double";
    if (const auto *m = s.find("model"))
      if (const auto *n = m->find("name")) name = n->as_string();
    if (const auto *po = s.find("plan_options"))
      if (const auto *p = po->find("precision")) precision = p->as_string();
    if (name == "tungsten" && precision == "double")
      return pfc::ui::App<pfc::TungstenHIP<double>>(s).main();
    if (name == "tungsten" && precision == "float")
      return pfc::ui::App<pfc::TungstenHIP<float>>(s).main();
    if (name == "pfc" && precision == "double") return pfc::ui::App<pfc::PFCModelHIP>(s).main();
    std::cerr << "unknown model '" << name << "' / precision '" << precision
              << "' (tungsten: double|float, pfc: double)\n";
    return EXIT_FAILURE;
  } catch (const std::exception &e) {
    std::cerr << e.what() << '\n';
    return EXIT_FAILURE;
  }
}
