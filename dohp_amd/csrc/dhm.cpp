// DHM field files (HDF5) of block function spaces and their closure vectors: the on-disk format of dohp's viewer
// (src/viewer/dhm.c, src/fs/impls/cont/contviewdhm.c), written and read with the HDF5 C library.  Host code only; a
// library of its own so that the operator library does not depend on HDF5.
#include <hdf5.h>

#include <cmath>
#include <cstdarg>
#include <cstddef>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dohp_dhm.h"

namespace {

thread_local std::string g_err;

dErr fail(const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return 1;
}

#define DHM_H5(expr, what)                                     \
  do {                                                         \
    if ((expr) < 0) return fail("HDF5 call failed: %s", what); \
  } while (0)
#define DHM_CALL(expr)     \
  do {                     \
    dErr e_ = (expr);      \
    if (e_) return e_;     \
  } while (0)

// In-memory records of the committed compound types (the members of the reference's dht_* structs,
// include/dohpviewerdhm.h:63-97).
struct Units {
  char *dimensions;
  double scale;
};
struct Field {
  char *name;
  Units units;
};
struct RealWithUnits {
  double value;
  Units units;
};
struct FSRec {
  char *degree, *global_offset, *partition, *ordered_subdomain, *bstatus;
  hobj_ref_t mesh;
  double time;
  int internal_state, number_of_subdomains;
  double boundingbox[3][2];
  hvl_t fields;
};
struct VecRec {
  hdset_reg_ref_t fs;
  double time;
  Units units;
  int internal_state;
};

// Opens a link if it exists (dH5Gopen/dH5Dopen, dhm.c:769-806: no HDF5 error stack for a missing name).
hid_t open_group(hid_t loc, const char *name)
{
  return H5Lexists(loc, name, H5P_DEFAULT) > 0 ? H5Gopen(loc, name, H5P_DEFAULT) : -1;
}
hid_t open_dataset(hid_t loc, const char *name)
{
  return H5Lexists(loc, name, H5P_DEFAULT) > 0 ? H5Dopen(loc, name, H5P_DEFAULT) : -1;
}

}  // namespace

struct _p_dohpDHM {
  int mode = -1;
  hid_t file = -1, root = -1, mesh = -1, fs = -1, step = -1, type = -1, cur = -1;
  hid_t t_fstring = -1, t_mstring = -1, t_units = -1, t_bbox = -1, t_time = -1, t_field = -1, t_fs = -1, t_vec = -1;
  hid_t s_scalar = -1;
  double time = 0.0;
  std::string timeunits = "s";
  double timescale = 1.0;
  int stepnumber = -1, totalsteps = 0;
  int state = 0;  // internal state counter written with every record

  // Types, committed to /dohp/type when writing (dViewerDHMGetStringTypes and friends, dhm.c:24-103,142-210).
  dErr commit(const char *name, hid_t t)
  {
    if (mode != dohpDHM_READ && H5Lexists(type, name, H5P_DEFAULT) <= 0)
      DHM_H5(H5Tcommit2(type, name, t, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "H5Tcommit");
    return 0;
  }
  dErr types()
  {
    if (t_vec >= 0) return 0;
    t_fstring = H5Tcopy(H5T_C_S1);
    DHM_H5(H5Tset_size(t_fstring, H5T_VARIABLE), "H5Tset_size");
    DHM_CALL(commit("string", t_fstring));
    t_mstring = H5Tcopy(H5T_C_S1);
    DHM_H5(H5Tset_size(t_mstring, H5T_VARIABLE), "H5Tset_size");
    t_units = H5Tcreate(H5T_COMPOUND, sizeof(Units));
    H5Tinsert(t_units, "dimensions", offsetof(Units, dimensions), t_fstring);
    H5Tinsert(t_units, "scale", offsetof(Units, scale), H5T_NATIVE_DOUBLE);
    DHM_CALL(commit("Units_type", t_units));
    const hsize_t bbdims[2] = {3, 2};
    t_bbox = H5Tarray_create2(H5T_NATIVE_DOUBLE, 2, bbdims);
    DHM_CALL(commit("BoundingBox_type", t_bbox));
    t_time = H5Tcreate(H5T_COMPOUND, sizeof(RealWithUnits));
    H5Tinsert(t_time, "value", offsetof(RealWithUnits, value), H5T_NATIVE_DOUBLE);
    H5Tinsert(t_time, "units", offsetof(RealWithUnits, units), t_units);
    DHM_CALL(commit("Time_type", t_time));
    t_field = H5Tcreate(H5T_COMPOUND, sizeof(Field));
    H5Tinsert(t_field, "name", offsetof(Field, name), t_fstring);
    H5Tinsert(t_field, "units", offsetof(Field, units), t_units);
    DHM_CALL(commit("Field_type", t_field));
    hid_t fields = H5Tvlen_create(t_field);
    t_fs = H5Tcreate(H5T_COMPOUND, sizeof(FSRec));
    H5Tinsert(t_fs, "degree", offsetof(FSRec, degree), t_fstring);
    H5Tinsert(t_fs, "global_offset", offsetof(FSRec, global_offset), t_fstring);
    H5Tinsert(t_fs, "partition", offsetof(FSRec, partition), t_fstring);
    H5Tinsert(t_fs, "ordered_subdomain", offsetof(FSRec, ordered_subdomain), t_fstring);
    H5Tinsert(t_fs, "bstatus", offsetof(FSRec, bstatus), t_fstring);
    H5Tinsert(t_fs, "mesh", offsetof(FSRec, mesh), H5T_STD_REF_OBJ);
    H5Tinsert(t_fs, "time", offsetof(FSRec, time), H5T_NATIVE_DOUBLE);
    H5Tinsert(t_fs, "internal_state", offsetof(FSRec, internal_state), H5T_NATIVE_INT);
    H5Tinsert(t_fs, "number_of_subdomains", offsetof(FSRec, number_of_subdomains), H5T_NATIVE_INT);
    H5Tinsert(t_fs, "boundingbox", offsetof(FSRec, boundingbox), t_bbox);
    H5Tinsert(t_fs, "fields", offsetof(FSRec, fields), fields);
    H5Tclose(fields);
    DHM_CALL(commit("FS_type", t_fs));
    t_vec = H5Tcreate(H5T_COMPOUND, sizeof(VecRec));
    H5Tinsert(t_vec, "fs", offsetof(VecRec, fs), H5T_STD_REF_DSETREG);
    H5Tinsert(t_vec, "time", offsetof(VecRec, time), H5T_NATIVE_DOUBLE);
    H5Tinsert(t_vec, "units", offsetof(VecRec, units), t_units);
    H5Tinsert(t_vec, "internal_state", offsetof(VecRec, internal_state), H5T_NATIVE_INT);
    DHM_CALL(commit("Vec_type", t_vec));
    return 0;
  }

  // Current step group: created when writing, opened when reading (dViewerDHMGetStep, dhm.c:404-437).
  dErr current(hid_t *out)
  {
    if (cur < 0) {
      if (stepnumber < 0) {
        DHM_CALL(set_time(0.0));
      }
      char name[16];
      snprintf(name, sizeof name, "%03d", stepnumber);
      DHM_CALL(types());
      if (mode == dohpDHM_READ || H5Lexists(step, name, H5P_DEFAULT) > 0) {
        cur = open_group(step, name);
        if (cur < 0) return fail("DHM: no step %s", name);
        hid_t att = H5Aopen(cur, "time", H5P_DEFAULT);
        DHM_H5(att, "H5Aopen(time)");
        RealWithUnits t;
        DHM_H5(H5Aread(att, t_time, &t), "H5Aread(time)");
        time = t.value;
        timeunits = t.units.dimensions ? t.units.dimensions : "";
        timescale = t.units.scale;
        hid_t sp = H5Aget_space(att);
        H5Dvlen_reclaim(t_time, sp, H5P_DEFAULT, &t);
        H5Sclose(sp);
        H5Aclose(att);
      } else {
        cur = H5Gcreate2(step, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        DHM_H5(cur, "H5Gcreate(step)");
        hid_t att = H5Acreate2(cur, "time", t_time, s_scalar, H5P_DEFAULT, H5P_DEFAULT);
        DHM_H5(att, "H5Acreate(time)");
        RealWithUnits t{time, {(char *)timeunits.c_str(), timescale}};
        DHM_H5(H5Awrite(att, t_time, &t), "H5Awrite(time)");
        H5Aclose(att);
      }
    }
    *out = cur;
    return 0;
  }
  void invalidate()
  {
    if (cur >= 0) H5Gclose(cur);
    cur = -1;
  }
  // dViewerDHMSetTime_DHM (dhm.c:377-390): a new step when the time changes (or at the first call).
  dErr set_time(double t)
  {
    if (!totalsteps || t != time) {
      stepnumber++;
      totalsteps++;
      invalidate();
    }
    time = t;
    return 0;
  }
  ~_p_dohpDHM()
  {
    invalidate();
    for (hid_t t : {t_fstring, t_mstring, t_units, t_bbox, t_time, t_field, t_fs, t_vec})
      if (t >= 0) H5Tclose(t);
    if (s_scalar >= 0) H5Sclose(s_scalar);
    for (hid_t g : {mesh, fs, step, type, root})
      if (g >= 0) H5Gclose(g);
    if (file >= 0) H5Fclose(file);
  }
};

extern "C" void dohpDHMGetLastError(char *buf, size_t len)
{
  if (!buf || !len) return;
  std::snprintf(buf, len, "%s", g_err.c_str());
}

extern "C" dErr dohpDHMOpen(const char *path, int mode, dohpDHM *out)
{
  if (!path || !out) return fail("null argument");
  if (mode < dohpDHM_READ || mode > dohpDHM_APPEND) return fail("DHM mode must be READ, WRITE or APPEND");
  *out = nullptr;
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);  // errors are returned, not printed
  _p_dohpDHM *v = new _p_dohpDHM;
  v->mode = mode;
  if (mode == dohpDHM_WRITE) v->file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  else v->file = H5Fopen(path, mode == dohpDHM_READ ? H5F_ACC_RDONLY : H5F_ACC_RDWR, H5P_DEFAULT);
  if (v->file < 0) {
    delete v;
    return fail("DHM: cannot %s \"%s\"", mode == dohpDHM_WRITE ? "create" : "open", path);
  }
  if (mode == dohpDHM_WRITE) {
    v->root = H5Gcreate2(v->file, "dohp", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    v->mesh = H5Gcreate2(v->root, "mesh", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    v->fs = H5Gcreate2(v->root, "fs", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    v->step = H5Gcreate2(v->root, "step", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    v->type = H5Gcreate2(v->root, "type", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  } else {
    v->root = open_group(v->file, "dohp");
    if (v->root >= 0) {
      v->mesh = open_group(v->root, "mesh");
      v->fs = open_group(v->root, "fs");
      v->step = open_group(v->root, "step");
      v->type = open_group(v->root, "type");
    }
  }
  if (v->root < 0 || v->mesh < 0 || v->fs < 0 || v->step < 0 || v->type < 0) {
    delete v;
    return fail("DHM: \"%s\" lacks the /dohp/{mesh,fs,step,type} groups", path);
  }
  v->s_scalar = H5Screate(H5S_SCALAR);
  if (mode != dohpDHM_WRITE) {  // steps already present
    H5G_info_t info;
    if (H5Gget_info(v->step, &info) >= 0) v->totalsteps = (int)info.nlinks;
    if (mode == dohpDHM_APPEND) v->stepnumber = v->totalsteps - 1;
    else if (v->totalsteps) v->stepnumber = 0;
  }
  *out = v;
  return 0;
}

extern "C" dErr dohpDHMClose(dohpDHM *v)
{
  if (v && *v) {
    H5Fflush((*v)->file, H5F_SCOPE_LOCAL);
    delete *v;
    *v = nullptr;
  }
  return 0;
}

extern "C" dErr dohpDHMSetTime(dohpDHM v, double time)
{
  if (!v) return fail("null viewer");
  if (v->mode == dohpDHM_READ) return fail("DHM: the file is open for reading (select steps with dohpDHMSetTimeStep)");
  return v->set_time(time);
}

extern "C" dErr dohpDHMSetTimeUnits(dohpDHM v, const char *units, double scale)
{
  if (!v || !units) return fail("null argument");
  v->timeunits = units;
  v->timescale = scale;
  return 0;
}

extern "C" dErr dohpDHMGetSteps(dohpDHM v, int *nsteps, double *times, int cap)
{
  if (!v || !nsteps) return fail("null argument");
  DHM_CALL(v->types());
  int k = 0;
  for (int s = 0; s < v->totalsteps; s++) {
    char name[16];
    snprintf(name, sizeof name, "%03d", s);
    hid_t g = open_group(v->step, name);
    if (g < 0) continue;
    hid_t att = H5Aopen(g, "time", H5P_DEFAULT);
    if (att >= 0) {
      RealWithUnits t;
      if (H5Aread(att, v->t_time, &t) >= 0) {
        if (times && k < cap) times[k] = t.value;
        hid_t sp = H5Aget_space(att);
        H5Dvlen_reclaim(v->t_time, sp, H5P_DEFAULT, &t);
        H5Sclose(sp);
        k++;
      }
      H5Aclose(att);
    }
    H5Gclose(g);
  }
  *nsteps = k;
  return 0;
}

extern "C" dErr dohpDHMSetTimeStep(dohpDHM v, int s)
{
  if (!v) return fail("null viewer");
  if (s < 0 || s >= v->totalsteps) return fail("Step %d out of range [0 .. %d]", s, v->totalsteps);  // dhm.c:493
  if (s != v->stepnumber) v->invalidate();
  v->stepnumber = s;
  return 0;
}

extern "C" dErr dohpDHMWriteBlockSpace(dohpDHM v, const char *fsname, int M0, int M1, int M2, int degree,
                                       int dirichlet_mask, const double *vcoords, int bs,
                                       const char *const *fieldnames)
{
  if (!v || !fsname || !vcoords) return fail("null argument");
  if (v->mode == dohpDHM_READ) return fail("DHM: the file is open for reading");
  if (M0 < 1 || M1 < 1 || M2 < 1 || degree < 1 || bs < 1) return fail("DHM: bad block space");
  DHM_CALL(v->types());
  if (H5Lexists(v->fs, fsname, H5P_DEFAULT) > 0) return 0;
  v->state++;
  // mesh record: /dohp/mesh/<fsname>-mesh/<state> (the reference: the MOAB file name, dhm.c:105-140)
  const std::string meshname = std::string(fsname) + "-mesh";
  hid_t mg = H5Gcreate2(v->mesh, meshname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  DHM_H5(mg, "H5Gcreate(mesh)");
  char state[16];
  snprintf(state, sizeof state, "%03d", v->state);
  char desc[96];
  snprintf(desc, sizeof desc, "block %d %d %d", M0, M1, M2);
  const char *dp = desc;
  hid_t md = H5Dcreate2(mg, state, v->t_fstring, v->s_scalar, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  DHM_H5(md, "H5Dcreate(mesh record)");
  DHM_H5(H5Dwrite(md, v->t_mstring, H5S_ALL, H5S_ALL, H5P_DEFAULT, &dp), "H5Dwrite(mesh record)");
  H5Dclose(md);
  const long nvert = (long)(M0 + 1) * (M1 + 1) * (M2 + 1);
  const hsize_t cdims[2] = {(hsize_t)nvert, 3};
  hid_t csp = H5Screate_simple(2, cdims, nullptr);
  hid_t cd = H5Dcreate2(mg, "coordinates", H5T_NATIVE_DOUBLE, csp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  DHM_H5(cd, "H5Dcreate(coordinates)");
  DHM_H5(H5Dwrite(cd, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, vcoords), "H5Dwrite(coordinates)");
  H5Dclose(cd);
  H5Sclose(csp);
  FSRec r{};
  DHM_H5(H5Rcreate(&r.mesh, mg, state, H5R_OBJECT, -1), "H5Rcreate(mesh)");
  H5Gclose(mg);
  // the strings the reference fills with MOAB tag names hold the block space's layout
  std::string deg = std::to_string(degree), goff = "lexicographic (I,J,K) K fastest, owned nodes first",
              part = "0", ords = "owned nodes, then Dirichlet nodes", bst = "dirichlet_mask " + std::to_string(dirichlet_mask);
  r.degree = (char *)deg.c_str();
  r.global_offset = (char *)goff.c_str();
  r.partition = (char *)part.c_str();
  r.ordered_subdomain = (char *)ords.c_str();
  r.bstatus = (char *)bst.c_str();
  r.time = v->time;
  r.internal_state = v->state;
  r.number_of_subdomains = 1;
  for (int d = 0; d < 3; d++) {
    r.boundingbox[d][0] = INFINITY;
    r.boundingbox[d][1] = -INFINITY;
  }
  for (long i = 0; i < nvert; i++)
    for (int d = 0; d < 3; d++) {
      r.boundingbox[d][0] = std::fmin(r.boundingbox[d][0], vcoords[3 * i + d]);
      r.boundingbox[d][1] = std::fmax(r.boundingbox[d][1], vcoords[3 * i + d]);
    }
  std::vector<std::string> names(bs);
  std::vector<Field> fields(bs);
  std::string dim = "";
  for (int i = 0; i < bs; i++) {
    names[i] = fieldnames && fieldnames[i] ? fieldnames[i] : "u" + std::to_string(i);
    fields[i].name = (char *)names[i].c_str();
    fields[i].units = Units{(char *)dim.c_str(), 1.0};
  }
  r.fields.len = bs;
  r.fields.p = fields.data();
  // one-record dataset, chunked and extensible (dFSGetDHMLink, contviewdhm.c:16-56)
  const hsize_t one[1] = {1}, unl[1] = {H5S_UNLIMITED};
  hid_t sp = H5Screate_simple(1, one, unl);
  hid_t dcpl = H5Pcreate(H5P_DATASET_CREATE);
  H5Pset_chunk(dcpl, 1, one);
  hid_t d = H5Dcreate2(v->fs, fsname, v->t_fs, sp, H5P_DEFAULT, dcpl, H5P_DEFAULT);
  H5Pclose(dcpl);
  DHM_H5(d, "H5Dcreate(fs)");
  DHM_H5(H5Dwrite(d, v->t_fs, H5S_ALL, H5S_ALL, H5P_DEFAULT, &r), "H5Dwrite(fs)");
  H5Dclose(d);
  H5Sclose(sp);
  return 0;
}

extern "C" dErr dohpDHMReadBlockSpace(dohpDHM v, const char *fsname, int *M, int *degree, int *mask, int *bs,
                                      double *vcoords)
{
  if (!v || !fsname) return fail("null argument");
  DHM_CALL(v->types());
  hid_t d = open_dataset(v->fs, fsname);
  if (d < 0) return fail("Could not find FS by name \"%s\"", fsname);  // dhm.c:247
  FSRec r;
  hid_t fsp = H5Dget_space(d);
  const hsize_t first[1] = {0};
  H5Sselect_elements(fsp, H5S_SELECT_SET, 1, first);
  hid_t msp = H5Screate(H5S_SCALAR);
  herr_t h = H5Dread(d, v->t_fs, msp, fsp, H5P_DEFAULT, &r);
  H5Sclose(msp);
  if (h < 0) {
    H5Sclose(fsp);
    H5Dclose(d);
    return fail("H5Dread(fs)");
  }
  dErr err = 0;
  if (degree) *degree = std::atoi(r.degree);
  if (mask) *mask = r.bstatus && std::strncmp(r.bstatus, "dirichlet_mask ", 15) == 0 ? std::atoi(r.bstatus + 15) : 0;
  if (bs) *bs = (int)r.fields.len;
  hid_t mo = H5Rdereference2(d, H5P_DEFAULT, H5R_OBJECT, &r.mesh);
  if (mo < 0) err = fail("H5Rdereference(mesh)");
  int m[3] = {0, 0, 0};
  if (!err) {
    char *desc = nullptr;
    if (H5Dread(mo, v->t_mstring, H5S_ALL, H5S_ALL, H5P_DEFAULT, &desc) < 0 || !desc ||
        std::sscanf(desc, "block %d %d %d", &m[0], &m[1], &m[2]) != 3)
      err = fail("DHM: \"%s\" is not a block-space record", fsname);
    if (desc) {
      hid_t ssp = H5Screate(H5S_SCALAR);
      H5Dvlen_reclaim(v->t_mstring, ssp, H5P_DEFAULT, &desc);
      H5Sclose(ssp);
    }
  }
  if (!err && M)
    for (int i = 0; i < 3; i++) M[i] = m[i];
  if (!err && vcoords) {
    char path[512];
    const ssize_t len = H5Iget_name(mo, path, sizeof path);  // /dohp/mesh/<name>/<state>
    std::string dir(path, len > 0 ? (size_t)len : 0);
    dir = dir.substr(0, dir.rfind('/')) + "/coordinates";
    hid_t cd = H5Dopen2(v->file, dir.c_str(), H5P_DEFAULT);
    if (cd < 0 || H5Dread(cd, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, vcoords) < 0)
      err = fail("H5Dread(coordinates)");
    if (cd >= 0) H5Dclose(cd);
  }
  if (mo >= 0) H5Oclose(mo);
  H5Dvlen_reclaim(v->t_fs, fsp, H5P_DEFAULT, &r);
  H5Sclose(fsp);
  H5Dclose(d);
  return err;
}

extern "C" dErr dohpDHMWriteVec(dohpDHM v, const char *fsname, const char *vecname, long n, int bs, const double *x,
                                double unit_scale)
{
  if (!v || !fsname || !vecname || !x) return fail("null argument");
  if (v->mode == dohpDHM_READ) return fail("DHM: the file is open for reading");
  if (bs < 1 || n < 0 || n % bs) return fail("DHM: vector size %ld is not a multiple of the block size %d", n, bs);
  hid_t fsd = open_dataset(v->fs, fsname);
  if (fsd < 0) return fail("DHM: write the space \"%s\" before its vectors", fsname);
  hid_t step;
  dErr err = v->current(&step);
  if (err) {
    H5Dclose(fsd);
    return err;
  }
  v->state++;
  // [n/bs][bs] (dVecGetHDF5Hyperslab, contviewdhm.c:113-130; one subdomain)
  const hsize_t dims[2] = {(hsize_t)(n / bs), (hsize_t)bs};
  hid_t sp = H5Screate_simple(2, dims, nullptr);
  if (H5Lexists(step, vecname, H5P_DEFAULT) > 0) H5Ldelete(step, vecname, H5P_DEFAULT);
  hid_t d = H5Dcreate2(step, vecname, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  if (d < 0 || H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, x) < 0) err = fail("H5Dwrite(vector)");
  if (!err) {  // "meta" attribute (contviewdhm.c:175-199)
    VecRec m{};
    char fspath[512];
    H5Iget_name(fsd, fspath, sizeof fspath);
    hid_t fsp = H5Dget_space(fsd);
    const hsize_t first[1] = {0};
    H5Sselect_elements(fsp, H5S_SELECT_SET, 1, first);
    if (H5Rcreate(&m.fs, v->file, fspath, H5R_DATASET_REGION, fsp) < 0) err = fail("H5Rcreate(fs region)");
    H5Sclose(fsp);
    m.time = v->time;
    std::string none = "";
    m.units = Units{(char *)none.c_str(), unit_scale};
    m.internal_state = v->state;
    const hsize_t one[1] = {1};
    hid_t asp = H5Screate_simple(1, one, nullptr);
    hid_t a = err ? -1 : H5Acreate2(d, "meta", v->t_vec, asp, H5P_DEFAULT, H5P_DEFAULT);
    if (!err && (a < 0 || H5Awrite(a, v->t_vec, &m) < 0)) err = fail("H5Awrite(meta)");
    if (a >= 0) H5Aclose(a);
    H5Sclose(asp);
  }
  if (d >= 0) H5Dclose(d);
  H5Sclose(sp);
  H5Dclose(fsd);
  return err;
}

extern "C" dErr dohpDHMReadVec(dohpDHM v, const char *vecname, long n, int bs, double *x)
{
  if (!v || !vecname || !x) return fail("null argument");
  hid_t step;
  DHM_CALL(v->current(&step));
  hid_t d = open_dataset(step, vecname);
  if (d < 0) return fail("DHM: no vector \"%s\" in step %03d", vecname, v->stepnumber);
  hid_t sp = H5Dget_space(d);
  hsize_t dims[2] = {0, 0};
  const int nd = H5Sget_simple_extent_dims(sp, dims, nullptr);
  dErr err = 0;
  if (nd != 2 || (long)(dims[0] * dims[1]) != n || (long)dims[1] != bs)
    err = fail("DHM: vector \"%s\" is [%llu][%llu], expected %ld values with block size %d", vecname,
               (unsigned long long)dims[0], (unsigned long long)dims[1], n, bs);
  if (!err && H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, x) < 0) err = fail("H5Dread(vector)");
  H5Sclose(sp);
  H5Dclose(d);
  return err;
}
