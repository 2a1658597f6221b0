// Field output of the host layer: Outputs (src/io/outputs.hpp:15-134), a
// NamedFieldsMap of data collections written each step, and VTKDataCollection,
// a DataCollection that writes the registered ND fields as cell-centre vectors
// of a legacy-VTK structured grid (one file per cycle, readable by ParaView /
// VisIt) with the element attributes beside them.
#pragma once

#include "core.hpp"

#include <string>
#include <vector>

namespace hephaestus
{

// Physical value of an ND_p field (true-DOF vector on the host) at the centre
// of every element: u = J^-T u_hat, u_hat from the tensor basis at the
// reference centre, J from the element's vertices (or the box spacing).
// Returns 3 values per element, element order x fastest.
std::vector<double> EvaluateCellCenters(const StructuredHexMesh & mesh, int order, const std::vector<double> & dofs);

class DataCollection
{
public:
  explicit DataCollection(std::string prefix) : _prefix(std::move(prefix)) {}
  virtual ~DataCollection() = default;
  void RegisterField(const std::string & name, GridFunction * gf) { _fields.emplace_back(name, gf); }
  void SetCycle(int c) { _cycle = c; }
  void SetTime(double t) { _time = t; }
  const std::string & Prefix() const { return _prefix; }
  virtual void Save() = 0;

protected:
  std::string _prefix;
  std::vector<std::pair<std::string, GridFunction *>> _fields; // not owned
  int _cycle = 0;
  double _time = 0.0;
};

// <prefix>_<cycle>.vtk per Save()
class VTKDataCollection : public DataCollection
{
public:
  using DataCollection::DataCollection;
  void Save() override;
  std::string LastFile() const { return _last; }

private:
  std::string _last;
};

class Outputs : public NamedFieldsMap<DataCollection>
{
public:
  Outputs() = default;
  explicit Outputs(GridFunctions & gridfunctions) : _gridfunctions(&gridfunctions) {}
  // fields to write; empty = every grid function
  void SetOutputFieldNames(const std::vector<std::string> & names) { _output_field_names = names; }
  void Init(GridFunctions & gridfunctions)
  {
    _gridfunctions = &gridfunctions;
    Reset();
  }
  // register the output fields with every collection and write cycle 0
  void Reset();
  // next cycle at time t
  void Write(double t = 1.0);
  int Cycle() const { return _cycle; }

private:
  GridFunctions * _gridfunctions = nullptr;
  std::vector<std::string> _output_field_names;
  int _cycle = 0;
  void WriteOutputFields(double t);
};

} // namespace hephaestus
