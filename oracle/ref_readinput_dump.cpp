// TEST INFRASTRUCTURE ONLY.  Runs the reference's own reader,
// readInput() of kernel/readInputFile.cpp (compiled from its source where it
// lies, READINPUT_SRC), on a problem file and dumps every array it parsed to
// <dir>/<name>.bin (raw float64, or int32 for the integer arrays), and the
// 100-entry header to <dir>/ndims.bin.  Tests compare these with
// hdg_amd.binio's reader and writer.
#define main readinput_reference_main
#include READINPUT_SRC
#undef main

#include <cstdio>
#include <string>

static void dump_raw(const std::string& path, const void* p, size_t bytes) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  if (bytes) std::fwrite(p, 1, bytes, f);
  std::fclose(f);
}
static void dump(const std::string& dir, const char* name, const std::vector<double>& a) {
  dump_raw(dir + "/" + name + ".bin", a.data(), a.size() * sizeof(double));
}
static void dump(const std::string& dir, const char* name, const std::vector<int>& a) {
  dump_raw(dir + "/" + name + ".bin", a.data(), a.size() * sizeof(int));
}

int main(int argc, char** argv) {
  if (argc != 3) { std::fprintf(stderr, "usage: %s problem.bin outdir\n", argv[0]); return 1; }
  std::vector<int> ndims;
  meshstruct mesh;
  masterstruct master;
  appstruct app;
  intstruct intf;
  solstruct sol;
  readInput(argv[1], ndims, mesh, master, app, intf, sol);
  const std::string d = argv[2];
  dump(d, "ndims", ndims);
  dump(d, "UDG", sol.UDG); dump(d, "UH", sol.UH);
  dump(d, "dgnodes", mesh.dgnodes); dump(d, "elcon", mesh.elcon); dump(d, "bf", mesh.bf); dump(d, "f", mesh.f);
  dump(d, "t2f", mesh.t2f); dump(d, "f2f", mesh.f2f); dump(d, "perm", mesh.perm); dump(d, "permgeom", mesh.permgeom);
  dump(d, "plocvl", master.plocvl); dump(d, "gpvl", master.gpvl); dump(d, "gwvl", master.gwvl);
  dump(d, "plocfc", master.plocfc); dump(d, "gpfc", master.gpfc); dump(d, "gwfc", master.gwfc);
  dump(d, "shapvt", master.shapvt); dump(d, "shapvg", master.shapvg); dump(d, "shapvgdotshapvl", master.shapvgdotshapvl);
  dump(d, "shapft", master.shapft); dump(d, "shapfg", master.shapfg); dump(d, "shapfgdotshapfc", master.shapfgdotshapfc);
  dump(d, "shapmv", master.shapmv); dump(d, "shapmf", master.shapmf);
  dump(d, "bcm", app.bcm); dump(d, "bcs", app.bcs); dump(d, "bcd", app.bcd); dump(d, "bcv", app.bcv);
  dump(d, "dt", app.dt); dump(d, "param", app.param); dump(d, "flag", app.flag); dump(d, "factor", app.factor);
  return 0;
}
