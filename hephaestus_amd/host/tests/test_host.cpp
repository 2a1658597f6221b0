// Tests of the host C++ layer.
//   test_host cpu                         framework logic, no device needed
//   test_host gpu <dir> <n> <p> <steps> <dt> <precision> [Jacobi|AMS] [Coil|DivFree|DivFreeProj]
//   test_host av <dir> <n> <p> <steps> <dt>  AV formulation steps (a, da_dt, v per step)
//   test_host meshio <dir> <n> <p>            apply on a mesh read back from an MFEM mesh file
//                                         backward-Euler steps of the H(curl)
//                                         curl-curl system on the device; writes
//                                         u after each step to <dir>/u_<k>.bin
#include "hephaestus/equation_system.hpp"
#include "hephaestus/mesh_io.hpp"
#include "hephaestus/outputs.hpp"

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>

using namespace hephaestus;

static int g_fail = 0;
#define EXPECT(cond)                                                                                             \
  do                                                                                                             \
  {                                                                                                              \
    if (!(cond))                                                                                                 \
    {                                                                                                            \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                                     \
      g_fail++;                                                                                                  \
    }                                                                                                            \
  } while (0)

template <typename F>
static bool
throws(F f)
{
  try
  {
    f();
  }
  catch (const Error &)
  {
    return true;
  }
  return false;
}

static void
test_cpu()
{
  // InputParameters
  InputParameters p;
  p.SetParam("CoefficientName", std::string("sigma"));
  p.SetParam("MaxIter", 50);
  EXPECT(p.GetParam<std::string>("CoefficientName") == "sigma");
  EXPECT(p.GetParam<int>("MaxIter") == 50);
  EXPECT(p.GetOptionalParam<double>("Tolerance", 1e-9) == 1e-9);
  EXPECT(throws([&] { p.GetParam<std::string>("Missing"); }));
  EXPECT(throws([&] { p.GetParam<double>("MaxIter"); })); // wrong type
  // NamedFieldsMap
  NamedFieldsMap<ConstantCoefficient> m;
  m.Register("a", std::make_shared<ConstantCoefficient>(2.0));
  EXPECT(m.Has("a") && !m.Has("b") && m.Get("a")->constant == 2.0);
  EXPECT(throws([&] { m.Get("b"); }));
  // coefficients on a two-region mesh
  StructuredHexMesh mesh = StructuredHexMesh::TwoRegionCube(8);
  int inner = 0;
  for (int a : mesh.attributes)
    inner += (a == 2);
  EXPECT(inner == 4 * 4 * 4);
  Coefficients coefs;
  coefs._scalars.Register("sigma", std::make_shared<PWConstCoefficient>(std::vector<double>{1e-6, 1e4}));
  coefs._scalars.Register("nu", std::make_shared<ConstantCoefficient>(1.0));
  auto sv = coefs._scalars.Get("sigma")->ElementValues(mesh);
  int64_t e_in = (4 * 8 + 4) * 8 + 4, e_out = 0;
  EXPECT(sv[e_in] == 1e4 && sv[e_out] == 1e-6);
  ConstantCoefficient dt(0.25);
  TransformedCoefficient dtnu(&dt, coefs._scalars.Get("nu"), prodFunc);
  EXPECT(dtnu.Eval(mesh, 3) == 0.25);
  TransformedCoefficient frac(coefs._scalars.Get("nu"), &dt, fracFunc);
  EXPECT(frac.Eval(mesh, 3) == 4.0);
  // FE space sizes: N = 3 (np)(np+1)^2 on an n^3 cube
  FESpace fes(&mesh, "ND_3D_P2");
  EXPECT(fes.Order() == 2 && fes.GetTrueVSize() == 3 * 16 * 17 * 17);
  FESpace h1(&mesh, "H1_3D_P2");
  EXPECT(h1.IsH1() && h1.GetTrueVSize() == 17 * 17 * 17);
  EXPECT(throws([&] { FESpace bad(&mesh, "L2_3D_P1"); }));
  EXPECT(throws([&] { FESpace bad(&mesh, "ND_3D_P7"); }));
  // boundary conditions
  BCMap bcs;
  bcs.Register("pec_x", std::make_shared<VectorDirichletBC>("pec_x", HC_FACE_XMIN | HC_FACE_XMAX));
  bcs.Register("pec_yz", std::make_shared<VectorDirichletBC>("pec_yz", HC_FACE_ALL & ~(HC_FACE_XMIN | HC_FACE_XMAX)));
  EXPECT(bcs.EssentialFaces() == HC_FACE_ALL);
  // kernel registration of the curl-curl formulation
  InputParameters esp;
  esp.SetParam("HCurlVarName", std::string("h"));
  esp.SetParam("AlphaCoefName", std::string("nu"));
  esp.SetParam("BetaCoefName", std::string("sigma"));
  CurlCurlEquationSystem es(esp);
  es.AddKernels();
  EXPECT(es._test_var_names.size() == 1 && es._test_var_names[0] == "dh_dt");
  EXPECT(es._trial_var_names.size() == 1 && es._trial_var_names[0] == "h");
  EXPECT(es._trial_var_time_derivative_names.size() == 1);
  EXPECT(throws([&] {
    InputParameters bad;
    CurlCurlKernel k(bad); // CoefficientName missing
  }));
  // cell-centre evaluation: a constant field u on an affine (sheared) mesh
  // has DOFs (J e_c) . u, and every cell returns u
  {
    StructuredHexMesh am;
    am.nx = 3;
    am.ny = 2;
    am.nz = 2;
    const double A[3][3] = {{1.0, 0.2, 0.0}, {0.0, 0.5, 0.1}, {0.3, 0.0, 0.8}};
    for (int k = 0; k <= am.nz; k++)
      for (int j = 0; j <= am.ny; j++)
        for (int i = 0; i <= am.nx; i++)
        {
          const double r[3] = {(double)i / am.nx, (double)j / am.ny, (double)k / am.nz};
          for (int a = 0; a < 3; a++)
            am.vert.push_back(A[a][0] * r[0] + A[a][1] * r[1] + A[a][2] * r[2] + 0.1 * a);
        }
    const double u[3] = {1.0, -2.0, 0.5};
    const int n[3] = {am.nx, am.ny, am.nz};
    for (int p = 1; p <= 3; p++)
    {
      std::vector<double> dofs;
      for (int c = 0; c < 3; c++)
      {
        // J e_c = A e_c / n_c (reference edge of unit length)
        double t = 0.0;
        for (int a = 0; a < 3; a++)
          t += A[a][c] / n[c] * u[a];
        int64_t cnt = 1;
        for (int a = 0; a < 3; a++)
          cnt *= (a == c) ? (int64_t)n[a] * p : (int64_t)n[a] * p + 1;
        dofs.insert(dofs.end(), cnt, t);
      }
      const std::vector<double> cc = EvaluateCellCenters(am, p, dofs);
      double err = 0.0;
      for (int64_t e = 0; e < am.NumElements(); e++)
        for (int a = 0; a < 3; a++)
          err = std::max(err, std::fabs(cc[3 * e + a] - u[a]));
      EXPECT(err <= 1e-13);
    }
    EXPECT(throws([&] { EvaluateCellCenters(am, 2, std::vector<double>(5)); }));
    {
      // all vertices of the plane z = 0 moved onto z = 1/nz: the first layer of
      // elements has zero volume, and the evaluation refuses it
      StructuredHexMesh flat = am;
      for (int j = 0; j <= am.ny; j++)
        for (int i = 0; i <= am.nx; i++)
        {
          const size_t v0 = (size_t)j * (am.nx + 1) + i;
          const size_t v1 = v0 + (size_t)(am.nx + 1) * (am.ny + 1);
          for (int a = 0; a < 3; a++)
            flat.vert[3 * v0 + a] = am.vert[3 * v1 + a];
        }
      std::vector<double> dofs;
      for (int c = 0; c < 3; c++)
      {
        int64_t cnt = 1;
        for (int a = 0; a < 3; a++)
          cnt *= (a == c) ? (int64_t)n[a] : (int64_t)n[a] + 1;
        dofs.insert(dofs.end(), cnt, 1.0);
      }
      EXPECT(throws([&] { EvaluateCellCenters(flat, 1, dofs); }));
    }
  }
  // kernel registration of the AV formulation (av_formulation.cpp:150-196)
  {
    InputParameters avp;
    avp.SetParam("VectorPotentialName", std::string("a"));
    avp.SetParam("ScalarPotentialName", std::string("v"));
    avp.SetParam("AlphaCoefName", std::string("nu"));
    avp.SetParam("BetaCoefName", std::string("sigma"));
    AVEquationSystem av(avp);
    av.AddKernels();
    EXPECT(av._test_var_names.size() == 2 && av._test_var_names[0] == "da_dt" && av._test_var_names[1] == "v");
    EXPECT(av._trial_var_names.size() == 2);
    EXPECT(av._dtalpha_coef_name == "dt_nu" && av._neg_beta_coef_name == "negative_sigma");
  }
  // MFEM mesh v1.0 round trips (box with attributes; sheared anisotropic grid),
  // a file with permuted vertex ids, and refusal of unstructured meshes
  {
    const std::string tmp = "/tmp/hephaestus_mesh_io_test.mesh";
    StructuredHexMesh box = StructuredHexMesh::TwoRegionCube(4);
    WriteMFEMMesh(box, tmp);
    StructuredHexMesh rb = ReadMFEMMesh(tmp);
    EXPECT(rb.nx == 4 && rb.ny == 4 && rb.nz == 4 && rb.attributes == box.attributes);
    bool same = rb.vert.size() == 3 * 125;
    for (int k = 0; k <= 4 && same; k++)
      for (int j = 0; j <= 4; j++)
        for (int i = 0; i <= 4; i++)
        {
          const size_t v = ((size_t)k * 5 + j) * 5 + i;
          same = same && rb.vert[3 * v] == i / 4.0 && rb.vert[3 * v + 1] == j / 4.0 && rb.vert[3 * v + 2] == k / 4.0;
        }
    EXPECT(same);
    StructuredHexMesh sh;
    sh.nx = 3;
    sh.ny = 2;
    sh.nz = 5;
    for (int k = 0; k <= sh.nz; k++)
      for (int j = 0; j <= sh.ny; j++)
        for (int i = 0; i <= sh.nx; i++)
        {
          sh.vert.push_back(0.3 * i + 0.05 * j * k);
          sh.vert.push_back(0.5 * j + 0.01 * i);
          sh.vert.push_back(0.2 * k - 0.02 * i * j);
        }
    sh.attributes.resize(sh.NumElements());
    for (int64_t e = 0; e < sh.NumElements(); e++)
      sh.attributes[e] = 1 + (int)(e % 3);
    WriteMFEMMesh(sh, tmp);
    StructuredHexMesh rs = ReadMFEMMesh(tmp);
    EXPECT(rs.nx == 3 && rs.ny == 2 && rs.nz == 5 && rs.vert == sh.vert && rs.attributes == sh.attributes);
    // the same mesh with its vertex ids reversed in the file
    {
      std::vector<std::string> lines;
      {
        std::ifstream in(tmp);
        std::string l;
        while (std::getline(in, l))
          lines.push_back(l);
      }
      const int64_t nv = 4 * 3 * 6;
      std::ofstream w(tmp);
      for (size_t i = 0; i < lines.size(); i++)
      {
        w << lines[i] << "\n";
        if (lines[i] == "elements")
        {
          const int64_t ne = std::stoll(lines[++i]);
          w << lines[i] << "\n";
          for (int64_t e = 0; e < ne; e++)
          {
            std::istringstream es(lines[++i]);
            int a, g;
            es >> a >> g;
            w << a << " " << g;
            int64_t v;
            while (es >> v)
              w << " " << (nv - 1 - v);
            w << "\n";
          }
        }
        else if (lines[i] == "vertices")
        {
          w << lines[i + 1] << "\n" << lines[i + 2] << "\n"; // count, dimension
          for (int64_t v = nv - 1; v >= 0; v--)
            w << lines[i + 3 + v] << "\n";
          i += 2 + nv;
        }
      }
    }
    StructuredHexMesh rp = ReadMFEMMesh(tmp);
    EXPECT(rp.nx == 3 && rp.ny == 2 && rp.nz == 5 && rp.vert == sh.vert);
    {
      std::ofstream w(tmp);
      w << "MFEM mesh v1.0\ndimension\n3\nelements\n1\n1 4 0 1 2 3\nboundary\n0\nvertices\n4\n3\n"
           "0 0 0\n1 0 0\n0 1 0\n0 0 1\n";
    }
    EXPECT(throws([&] { ReadMFEMMesh(tmp); }));
    {
      // a non-numeric count is an Error, not an uncaught std::invalid_argument
      std::ofstream w(tmp);
      w << "MFEM mesh v1.0\ndimension\nthree\n";
    }
    EXPECT(throws([&] { ReadMFEMMesh(tmp); }));
    {
      // negative and huge section counts are Errors, not length_error / bad_alloc
      std::ofstream w(tmp);
      w << "MFEM mesh v1.0\ndimension\n3\nelements\n0\nboundary\n0\nvertices\n-8\n3\n";
    }
    EXPECT(throws([&] { ReadMFEMMesh(tmp); }));
    {
      std::ofstream w(tmp);
      w << "MFEM mesh v1.0\ndimension\n3\nelements\n-1\n";
    }
    EXPECT(throws([&] { ReadMFEMMesh(tmp); }));
    {
      std::ofstream w(tmp);
      w << "MFEM mesh v1.0\ndimension\n3\nvertices\n99999999999\n3\n";
    }
    EXPECT(throws([&] { ReadMFEMMesh(tmp); }));
    EXPECT(throws([&] { ReadMFEMMesh("/nonexistent/x.mesh"); }));
    std::remove(tmp.c_str());
  }
  // without a device, allocation fails with an Error (not an abort)
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    EXPECT(throws([&] { DeviceVector v(16); }));
  std::printf("cpu tests: %s\n", g_fail ? "FAILED" : "ok");
}

static void
write_vec(const std::string & path, const std::vector<double> & v)
{
  std::ofstream f(path, std::ios::binary);
  f.write((const char *)v.data(), sizeof(double) * v.size());
}

static void
test_gpu(const std::string & dir, int n, int order, int steps, double dt, const std::string & precision,
         const std::string & preconditioner, const std::string & source)
{
  StructuredHexMesh mesh = StructuredHexMesh::TwoRegionCube(n);
  FESpaces fespaces;
  fespaces.Register("HCurl", std::make_shared<FESpace>(&mesh, "ND_3D_P" + std::to_string(order)));
  GridFunctions gfs;
  gfs.Register("h", std::make_shared<GridFunction>(fespaces.Get("HCurl")));
  gfs.Register("dh_dt", std::make_shared<GridFunction>(fespaces.Get("HCurl")));
  Coefficients coefs;
  coefs._scalars.Register("nu", std::make_shared<PWConstCoefficient>(std::vector<double>{1.0, 0.5}));
  coefs._scalars.Register("sigma", std::make_shared<PWConstCoefficient>(std::vector<double>{1e-2, 1e2}));
  BCMap bcs;
  bcs.Register("pec", std::make_shared<VectorDirichletBC>("pec", HC_FACE_ALL));
  Sources sources;
  std::shared_ptr<DivFreeSource> dfs;
  if (source == "Coil")
    sources.Register("coil", std::make_shared<CoilSource>(11));
  else
  {
    dfs = std::make_shared<DivFreeSource>(11, source == "DivFreeProj");
    sources.Register("coil", dfs);
  }
  InputParameters esp;
  esp.SetParam("HCurlVarName", std::string("h"));
  esp.SetParam("AlphaCoefName", std::string("nu"));
  esp.SetParam("BetaCoefName", std::string("sigma"));
  CurlCurlEquationSystem es(esp);
  es.AddKernels();
  es.Init(gfs, fespaces, bcs, coefs);
  InputParameters solver;
  solver.SetParam("Tolerance", 1e-12);
  solver.SetParam("MaxIter", 20000);
  solver.SetParam("Precision", precision);
  solver.SetParam("Preconditioner", preconditioner);
  TimeDomainEquationSystemProblemOperator op(es, gfs, bcs, coefs, sources, solver);
  for (int k = 0; k < steps; k++)
  {
    const hc_solve_stats & st = op.Step(dt, "h");
    std::printf("step %d: iterations %d outer %d rel_res %.3e converged %d\n", k, st.iterations,
                st.outer_iterations, st.rel_res, st.converged);
    EXPECT(st.converged == 1);
    write_vec(dir + "/u_" + std::to_string(k) + ".bin", gfs.Get("h")->Download());
  }
  if (dfs)
  {
    // SubtractSource takes back what Apply added
    DeviceVector t(dfs->SourceField().Size());
    t.Fill(0.0);
    dfs->SubtractSource(&t);
    std::vector<double> th = t.Download();
    double s2 = 0.0;
    for (double v : th)
      s2 += v * v;
    EXPECT(s2 > 0.0);
  }
  // field output: cycle 0 at Reset, cycle 1 after the steps
  {
    Outputs outs(gfs);
    outs.Register("vtk", std::make_shared<VTKDataCollection>(dir + "/fields"));
    outs.SetOutputFieldNames({"h"});
    outs.Reset();
    outs.Write(dt * steps);
    EXPECT(outs.Cycle() == 1);
  }
  // dt change re-assembles (dt rule), an unchanged dt does not
  BilinearForm & blf = es.GetGradient();
  es.SetTimeStep(dt);
  blf.Assemble();
  EXPECT(blf.LastAssembleWork() == 0);
  es.SetTimeStep(dt * 0.5);
  EXPECT(blf.LastAssembleWork() == 1);
  std::printf("gpu tests: %s\n", g_fail ? "FAILED" : "ok");
}

// Backward-Euler steps of the AV system with the projected coil source; writes
// A, dA/dt and V after each step to <dir>/{a,k,v}_<k>.bin.
static void
test_av(const std::string & dir, int n, int order, int steps, double dt)
{
  StructuredHexMesh mesh = StructuredHexMesh::TwoRegionCube(n);
  FESpaces fespaces;
  fespaces.Register("HCurl", std::make_shared<FESpace>(&mesh, "ND_3D_P" + std::to_string(order)));
  fespaces.Register("H1", std::make_shared<FESpace>(&mesh, "H1_3D_P" + std::to_string(order)));
  GridFunctions gfs;
  gfs.Register("a", std::make_shared<GridFunction>(fespaces.Get("HCurl")));
  gfs.Register("da_dt", std::make_shared<GridFunction>(fespaces.Get("HCurl")));
  gfs.Register("v", std::make_shared<GridFunction>(fespaces.Get("H1")));
  Coefficients coefs;
  coefs._scalars.Register("nu", std::make_shared<PWConstCoefficient>(std::vector<double>{1.0, 0.5}));
  coefs._scalars.Register("sigma", std::make_shared<PWConstCoefficient>(std::vector<double>{1e-2, 1e2}));
  BCMap bcs;
  bcs.Register("pec", std::make_shared<VectorDirichletBC>("pec", HC_FACE_ALL));
  Sources sources;
  sources.Register("coil", std::make_shared<DivFreeSource>(11, true));
  InputParameters avp;
  avp.SetParam("VectorPotentialName", std::string("a"));
  avp.SetParam("ScalarPotentialName", std::string("v"));
  avp.SetParam("AlphaCoefName", std::string("nu"));
  avp.SetParam("BetaCoefName", std::string("sigma"));
  AVEquationSystem es(avp);
  es.AddKernels();
  es.Init(gfs, fespaces, bcs, coefs);
  for (int k = 0; k < steps; k++)
  {
    const hc_solve_stats st = es.Step(dt, gfs, bcs, sources, 1e-11, 20000);
    std::printf("av step %d: iterations %d rel_res %.3e converged %d\n", k, st.iterations, st.rel_res, st.converged);
    EXPECT(st.converged == 1);
    write_vec(dir + "/a_" + std::to_string(k) + ".bin", gfs.Get("a")->Download());
    write_vec(dir + "/k_" + std::to_string(k) + ".bin", gfs.Get("da_dt")->Download());
    write_vec(dir + "/v_" + std::to_string(k) + ".bin", gfs.Get("v")->Download());
  }
  std::printf("av tests: %s\n", g_fail ? "FAILED" : "ok");
}

// An operator on a mesh read back from an MFEM mesh file (trilinear-vertex
// path) applies as the one on the box mesh it was written from.
static void
test_meshio_gpu(const std::string & dir, int n, int order)
{
  StructuredHexMesh box = StructuredHexMesh::TwoRegionCube(n);
  WriteMFEMMesh(box, dir + "/cube.mesh");
  StructuredHexMesh rd = ReadMFEMMesh(dir + "/cube.mesh");
  Coefficients coefs;
  coefs._scalars.Register("sigma", std::make_shared<PWConstCoefficient>(std::vector<double>{1e-2, 1e2}));
  coefs._scalars.Register("nu", std::make_shared<ConstantCoefficient>(1.0));
  std::vector<double> y[2];
  const StructuredHexMesh * meshes[2] = {&box, &rd};
  for (int k = 0; k < 2; k++)
  {
    FESpace fes(meshes[k], "ND_3D_P" + std::to_string(order));
    BilinearForm blf(&fes);
    blf.AddDomainIntegrator(CurlCurlIntegrator{coefs._scalars.Get("nu")});
    blf.AddDomainIntegrator(VectorFEMassIntegrator{coefs._scalars.Get("sigma")});
    blf.SetEssentialFaces(HC_FACE_ALL);
    blf.Assemble();
    DeviceVector x(fes.GetTrueVSize()), yy(fes.GetTrueVSize());
    CheckHc(hc_fill_uniform(blf.Handle(), 5, x.Data(), x.Size(), nullptr), "meshio");
    blf.Mult(x, yy);
    y[k] = yy.Download();
  }
  double d2 = 0.0, n2 = 0.0;
  for (size_t i = 0; i < y[0].size(); i++)
  {
    d2 += (y[1][i] - y[0][i]) * (y[1][i] - y[0][i]);
    n2 += y[0][i] * y[0][i];
  }
  std::printf("meshio: relative difference %.3e\n", std::sqrt(d2 / n2));
  EXPECT(y[0].size() == y[1].size() && d2 <= 1e-24 * n2);
  std::printf("meshio tests: %s\n", g_fail ? "FAILED" : "ok");
}

int
main(int argc, char ** argv)
{
  if (argc < 2)
  {
    std::fprintf(stderr, "usage: test_host cpu | gpu <dir> <n> <p> <steps> <dt> <precision> [Jacobi|AMS] "
                         "[Coil|DivFree|DivFreeProj]\n");
    return 2;
  }
  try
  {
    if (!std::strcmp(argv[1], "cpu"))
      test_cpu();
    else if (!std::strcmp(argv[1], "meshio") && argc >= 5)
      test_meshio_gpu(argv[2], std::atoi(argv[3]), std::atoi(argv[4]));
    else if (!std::strcmp(argv[1], "av") && argc >= 7)
      test_av(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atof(argv[6]));
    else if (!std::strcmp(argv[1], "gpu") && argc >= 8)
      test_gpu(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]), std::atof(argv[6]), argv[7],
               argc >= 9 ? argv[8] : "Jacobi", argc >= 10 ? argv[9] : "Coil");
    else
    {
      std::fprintf(stderr, "bad arguments\n");
      return 2;
    }
  }
  catch (const Error & e)
  {
    std::fprintf(stderr, "hephaestus::Error: %s\n", e.what());
    return 1;
  }
  return g_fail ? 1 : 0;
}
