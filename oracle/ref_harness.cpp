// Test infrastructure only: a minimal CPython binding that drives the
// reference CEUpdater (cpp/src/ce_updater.cpp, compiled from the reference
// tree by oracle/Makefile into oracle/_ref/).  It exposes the same methods the
// reference's own Cython wrapper exposes (cemc/cpp_ext/pyce_updater.pyx:5-60)
// so fixture scripts can call the real updater exactly as cemc.CE does
// (cemc/ce_calculator.py:199-210, 345-364).  Nothing in the product links or
// imports this module.
#include <Python.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "ce_updater.hpp"

namespace {

struct RefUpdater {
  PyObject_HEAD
  CEUpdater* upd;
  PyObject* keep;  // tuple (atoms, bc, cf, eci): kept alive like pyce_updater.pyx:20-24
};

// Maps C++ exceptions the way Cython's `except +` does: invalid_argument ->
// ValueError, other std::exception -> RuntimeError (cemc/cpp_ext/ce_updater.pxd).
#define GUARD(stmt)                                              \
  try {                                                          \
    stmt;                                                        \
  } catch (const std::invalid_argument& e) {                     \
    PyErr_SetString(PyExc_ValueError, e.what());                 \
    return nullptr;                                              \
  } catch (const std::exception& e) {                            \
    PyErr_SetString(PyExc_RuntimeError, e.what());               \
    return nullptr;                                              \
  }

PyObject* ru_new(PyTypeObject* type, PyObject*, PyObject*) {
  RefUpdater* self = reinterpret_cast<RefUpdater*>(type->tp_alloc(type, 0));
  if (self) {
    self->upd = new CEUpdater();
    self->keep = nullptr;
  }
  return reinterpret_cast<PyObject*>(self);
}

void ru_dealloc(RefUpdater* self) {
  delete self->upd;
  Py_XDECREF(self->keep);
  Py_TYPE(self)->tp_free(reinterpret_cast<PyObject*>(self));
}

int ru_init(RefUpdater* self, PyObject* args, PyObject*) {
  PyObject *atoms, *bc, *cf, *eci;
  if (!PyArg_ParseTuple(args, "OOOO", &atoms, &bc, &cf, &eci)) return -1;
  self->keep = PyTuple_Pack(4, atoms, bc, cf, eci);
  // ~CEUpdater releases one reference to atoms (ce_updater.cpp:22) that init
  // never took; take it here so the count stays balanced.
  Py_INCREF(atoms);
  try {
    self->upd->init(atoms, bc, cf, eci);
  } catch (const std::invalid_argument& e) {
    PyErr_SetString(PyExc_ValueError, e.what());
    return -1;
  } catch (const std::exception& e) {
    PyErr_SetString(PyExc_RuntimeError, e.what());
    return -1;
  }
  return 0;
}

PyObject* ru_clear_history(RefUpdater* self, PyObject*) {
  GUARD(self->upd->clear_history());
  Py_RETURN_NONE;
}

PyObject* ru_undo_changes(RefUpdater* self, PyObject*) {
  GUARD(self->upd->undo_changes());
  Py_RETURN_NONE;
}

PyObject* ru_update_cf(RefUpdater* self, PyObject* change) {
  GUARD(self->upd->update_cf(change));
  Py_RETURN_NONE;
}

PyObject* ru_calculate(RefUpdater* self, PyObject* changes) {
  double e = 0.0;
  GUARD(e = self->upd->calculate(changes));
  return PyFloat_FromDouble(e);
}

PyObject* ru_get_energy(RefUpdater* self, PyObject*) {
  double e = 0.0;
  GUARD(e = self->upd->get_energy());
  return PyFloat_FromDouble(e);
}

PyObject* ru_get_cf(RefUpdater* self, PyObject*) {
  PyObject* out = nullptr;
  GUARD(out = self->upd->get_cf());
  return out;
}

PyObject* ru_set_ecis(RefUpdater* self, PyObject* ecis) {
  GUARD(self->upd->set_ecis(ecis));
  Py_RETURN_NONE;
}

PyObject* ru_get_singlets(RefUpdater* self, PyObject*) {
  PyObject* out = nullptr;
  GUARD(out = self->upd->get_singlets());
  return out;
}

PyObject* ru_get_symbols(RefUpdater* self, PyObject*) {
  const std::vector<std::string>& s = self->upd->get_symbols();
  PyObject* out = PyList_New(static_cast<Py_ssize_t>(s.size()));
  for (size_t i = 0; i < s.size(); ++i)
    PyList_SET_ITEM(out, static_cast<Py_ssize_t>(i), PyUnicode_FromString(s[i].c_str()));
  return out;
}

PyObject* ru_set_num_threads(RefUpdater* self, PyObject* n) {
  long v = PyLong_AsLong(n);
  if (v <= 0) {
    PyErr_SetString(PyExc_ValueError, "num_threads must be positive");
    return nullptr;
  }
  self->upd->set_num_threads(static_cast<unsigned int>(v));
  Py_RETURN_NONE;
}

PyMethodDef ru_methods[] = {
    {"clear_history", (PyCFunction)ru_clear_history, METH_NOARGS, nullptr},
    {"undo_changes", (PyCFunction)ru_undo_changes, METH_NOARGS, nullptr},
    {"update_cf", (PyCFunction)ru_update_cf, METH_O, nullptr},
    {"calculate", (PyCFunction)ru_calculate, METH_O, nullptr},
    {"get_energy", (PyCFunction)ru_get_energy, METH_NOARGS, nullptr},
    {"get_cf", (PyCFunction)ru_get_cf, METH_NOARGS, nullptr},
    {"set_ecis", (PyCFunction)ru_set_ecis, METH_O, nullptr},
    {"get_singlets", (PyCFunction)ru_get_singlets, METH_NOARGS, nullptr},
    {"get_symbols", (PyCFunction)ru_get_symbols, METH_NOARGS, nullptr},
    {"set_num_threads", (PyCFunction)ru_set_num_threads, METH_O, nullptr},
    {nullptr, nullptr, 0, nullptr}};

PyTypeObject RefUpdaterType = {PyVarObject_HEAD_INIT(nullptr, 0)};

PyModuleDef ref_module = {PyModuleDef_HEAD_INIT, "cemc_ref",
                          "Reference CEUpdater driven for fixture generation", -1,
                          nullptr};

}  // namespace

PyMODINIT_FUNC PyInit_cemc_ref(void) {
  RefUpdaterType.tp_name = "cemc_ref.RefUpdater";
  RefUpdaterType.tp_basicsize = sizeof(RefUpdater);
  RefUpdaterType.tp_flags = Py_TPFLAGS_DEFAULT;
  RefUpdaterType.tp_new = ru_new;
  RefUpdaterType.tp_init = (initproc)ru_init;
  RefUpdaterType.tp_dealloc = (destructor)ru_dealloc;
  RefUpdaterType.tp_methods = ru_methods;
  if (PyType_Ready(&RefUpdaterType) < 0) return nullptr;
  PyObject* m = PyModule_Create(&ref_module);
  if (!m) return nullptr;
  Py_INCREF(&RefUpdaterType);
  PyModule_AddObject(m, "RefUpdater", reinterpret_cast<PyObject*>(&RefUpdaterType));
  return m;
}
