/*
 * Dump the numeric content of a JLD2 (HDF5) file as raw little-endian
 * arrays plus one JSON line per array on stdout.
 *
 * Used only to turn the reference's committed simulation outputs
 * (simulation/{test_test_D2_X32,hb_test_D2_X32}/*.jld2) into small test
 * fixtures.  JLD2 stores Julia structs as HDF5 compounds whose fields are
 * object references, and Vector{Array} as datasets of object references;
 * both are followed recursively.  Complex numbers are the compound {re, im}.
 *
 * Build:  gcc -O1 -o jld2_dump jld2_dump.c -I<hdf5>/include -L<hdf5>/lib -lhdf5
 * Usage:  jld2_dump file.jld2 outdir
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static const char *g_outdir;
static int g_first = 1;

static void emit(const char *name, const char *dtype, int rank, const hsize_t *dims,
                 const void *data, size_t nbytes) {
    char path[4096];
    snprintf(path, sizeof path, "%s/%s.bin", g_outdir, name);
    FILE *f = fopen(path, "wb");
    if (!f) { perror(path); exit(1); }
    fwrite(data, 1, nbytes, f);
    fclose(f);
    printf("{\"name\": \"%s\", \"dtype\": \"%s\", \"shape\": [", name, dtype);
    for (int i = 0; i < rank; ++i) printf("%s%llu", i ? ", " : "", (unsigned long long)dims[i]);
    printf("], \"file\": \"%s.bin\"}\n", name);
    (void)g_first;
}

static int is_complex_type(hid_t t) {
    if (H5Tget_class(t) != H5T_COMPOUND || H5Tget_nmembers(t) != 2) return 0;
    char *a = H5Tget_member_name(t, 0), *b = H5Tget_member_name(t, 1);
    int ok = a && b && !strcmp(a, "re") && !strcmp(b, "im");
    H5free_memory(a); H5free_memory(b);
    if (!ok) return 0;
    hid_t m0 = H5Tget_member_type(t, 0);
    ok = H5Tget_class(m0) == H5T_FLOAT && H5Tget_size(m0) == 8;
    H5Tclose(m0);
    return ok;
}

static void dump_dataset(hid_t dset, const char *name, int depth);

static void follow_ref(hid_t loc, hobj_ref_t *ref, const char *name, int depth) {
    if (depth > 12) return;
    hid_t obj = H5Rdereference2(loc, H5P_DEFAULT, H5R_OBJECT, ref);
    if (obj < 0) return;
    if (H5Iget_type(obj) == H5I_DATASET) dump_dataset(obj, name, depth + 1);
    H5Oclose(obj);
}

static void dump_dataset(hid_t dset, const char *name, int depth) {
    hid_t ftype = H5Dget_type(dset);
    hid_t space = H5Dget_space(dset);
    int rank = H5Sget_simple_extent_ndims(space);
    hsize_t dims[16] = {0};
    if (rank > 0) H5Sget_simple_extent_dims(space, dims, NULL);
    hssize_t n = H5Sget_simple_extent_npoints(space);
    H5T_class_t cls = H5Tget_class(ftype);

    if (cls == H5T_FLOAT && H5Tget_size(ftype) == 8) {
        double *buf = malloc(sizeof(double) * (n ? n : 1));
        H5Dread(dset, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
        emit(name, "f64", rank, dims, buf, sizeof(double) * n);
        free(buf);
    } else if (cls == H5T_INTEGER && H5Tget_size(ftype) == 8) {
        long long *buf = malloc(sizeof(long long) * (n ? n : 1));
        H5Dread(dset, H5T_NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
        emit(name, "i64", rank, dims, buf, sizeof(long long) * n);
        free(buf);
    } else if (is_complex_type(ftype)) {
        hid_t mt = H5Tcreate(H5T_COMPOUND, 16);
        H5Tinsert(mt, "re", 0, H5T_NATIVE_DOUBLE);
        H5Tinsert(mt, "im", 8, H5T_NATIVE_DOUBLE);
        double *buf = malloc(16 * (n ? n : 1));
        H5Dread(dset, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
        emit(name, "c128", rank, dims, buf, 16 * n);
        free(buf);
        H5Tclose(mt);
    } else if (cls == H5T_REFERENCE) {
        hobj_ref_t *refs = malloc(sizeof(hobj_ref_t) * (n ? n : 1));
        H5Dread(dset, H5T_STD_REF_OBJ, H5S_ALL, H5S_ALL, H5P_DEFAULT, refs);
        for (hssize_t i = 0; i < n; ++i) {
            char sub[2048];
            snprintf(sub, sizeof sub, "%s.%lld", name, (long long)i);
            follow_ref(dset, &refs[i], sub, depth);
        }
        free(refs);
    } else if (cls == H5T_COMPOUND && n == 1) {
        int nm = H5Tget_nmembers(ftype);
        for (int i = 0; i < nm; ++i) {
            char *mname = H5Tget_member_name(ftype, i);
            hid_t mtype = H5Tget_member_type(ftype, i);
            char sub[2048];
            snprintf(sub, sizeof sub, "%s.%s", name, mname);
            H5T_class_t mc = H5Tget_class(mtype);
            hid_t memt = -1;
            size_t sz = 0;
            if (mc == H5T_REFERENCE) { memt = H5T_STD_REF_OBJ; sz = sizeof(hobj_ref_t); }
            else if (mc == H5T_FLOAT && H5Tget_size(mtype) == 8) { memt = H5T_NATIVE_DOUBLE; sz = 8; }
            else if (mc == H5T_INTEGER && H5Tget_size(mtype) == 8) { memt = H5T_NATIVE_LLONG; sz = 8; }
            if (memt >= 0) {
                hid_t ct = H5Tcreate(H5T_COMPOUND, sz);
                H5Tinsert(ct, mname, 0, memt);
                unsigned char buf[64];
                if (H5Dread(dset, ct, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf) >= 0) {
                    if (mc == H5T_REFERENCE) follow_ref(dset, (hobj_ref_t *)buf, sub, depth);
                    else emit(sub, mc == H5T_FLOAT ? "f64" : "i64", 0, dims, buf, 8);
                }
                H5Tclose(ct);
            }
            H5Tclose(mtype);
            H5free_memory(mname);
        }
    }
    H5Sclose(space);
    H5Tclose(ftype);
}

static herr_t visit_top(hid_t g, const char *name, const H5L_info_t *info, void *op) {
    (void)info; (void)op;
    if (name[0] == '_') return 0;
    H5O_info_t oi;
    if (H5Oget_info_by_name1(g, name, &oi, H5P_DEFAULT) < 0) return 0;
    if (oi.type != H5O_TYPE_DATASET) return 0;
    hid_t d = H5Dopen2(g, name, H5P_DEFAULT);
    dump_dataset(d, name, 0);
    H5Dclose(d);
    return 0;
}

int main(int argc, char **argv) {
    if (argc != 3) { fprintf(stderr, "usage: %s file.jld2 outdir\n", argv[0]); return 2; }
    g_outdir = argv[2];
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) { fprintf(stderr, "cannot open %s\n", argv[1]); return 1; }
    H5Literate(f, H5_INDEX_NAME, H5_ITER_INC, NULL, visit_top, NULL);
    H5Fclose(f);
    return 0;
}
