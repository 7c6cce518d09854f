// JNI exports of class gpufitImFCS.GpufitImFCS over the C ABI.
//
// Symbol names, signatures and error behaviour follow the native methods
// declared in GpufitImFCS.java:323-538 and implemented in
// agpufitjni.cu:4595-5831: array methods read Java arrays with
// Get<Type>ArrayElements, write results with Set<Type>ArrayRegion and report
// failures with ThrowNew("java/lang/Exception"); `fit` works on direct NIO
// buffers and returns a status code with the message in getLastError.
#include "jni_min.h"

#include "../../include/imfcs.h"

#include <cstring>

namespace {

void throw_java(JNIEnv *env, const char *msg) {
    jclass ex = jni::FindClass(env, "java/lang/Exception");
    jni::ThrowNew(env, ex, msg);
}

// Reads the fields of GpufitImFCS.ACFParameters (GpufitImFCS.java:453-486).
void read_params(JNIEnv *env, jobject obj, imfcs_acf_params *p) {
    std::memset(p, 0, sizeof(*p));
    jclass c = jni::GetObjectClass(env, obj);
    auto gi = [&](const char *n) { return jni::GetIntField(env, obj, jni::GetFieldID(env, c, n, "I")); };
    auto gd = [&](const char *n) { return jni::GetDoubleField(env, obj, jni::GetFieldID(env, c, n, "D")); };
    auto gz = [&](const char *n) { return (int)jni::GetBooleanField(env, obj, jni::GetFieldID(env, c, n, "Z")); };
    p->width = gi("width");
    p->height = gi("height");
    p->win_star = gi("win_star");
    p->hin_star = gi("hin_star");
    p->w_temp = gi("w_temp");
    p->h_temp = gi("h_temp");
    p->pixbinX = gi("pixbinX");
    p->pixbinY = gi("pixbinY");
    p->binningX = gi("binningX");
    p->binningY = gi("binningY");
    p->firstframe = gi("firstframe");
    p->lastframe = gi("lastframe");
    p->framediff = p->lastframe - p->firstframe + 1;   // recomputed as the native side does
    p->cfXDistance = gi("cfXDistance");
    p->cfYDistance = gi("cfYDistance");
    p->correlatorp = gd("correlatorp");
    p->correlatorq = gd("correlatorq");
    p->frametime = gd("frametime");
    p->background = gi("background");
    p->mtab1 = gd("mtab1");
    p->mtabchanumminus1 = gd("mtabchanumminus1");
    p->sampchanumminus1 = gd("sampchanumminus1");
    p->chanum = gi("chanum");
    p->isNBcalculation = gz("isNBcalculation");
    p->bleachcorr_gpu = gz("bleachcorr_gpu");
    p->bleachcorr_order = gi("bleachcorr_order");
    p->nopit = gi("nopit");
    p->ave = gi("ave");
}

}  // namespace

extern "C" {

JNIEXPORT jint JNICALL Java_gpufitImFCS_GpufitImFCS_fit(
    JNIEnv *env, jclass, jint number_fits, jint number_points, jobject data_buffer, jobject weights_buffer,
    jint model_id, jobject initial_parameter_buffer, jfloat tolerance, jint max_number_iterations,
    jint num_valid_coefs, jobject parameters_to_fit_buffer, jint estimator_id, jint user_info_size,
    jobject user_info_buffer, jobject output_parameters_buffer, jobject output_states_buffer,
    jobject output_chi_squares_buffer, jobject output_number_iterations_buffer) {
    return imfcs_fit(number_fits, number_points, (const float *)jni::GetDirectBufferAddress(env, data_buffer),
                     (const float *)jni::GetDirectBufferAddress(env, weights_buffer), model_id,
                     (const float *)jni::GetDirectBufferAddress(env, initial_parameter_buffer), tolerance,
                     max_number_iterations, num_valid_coefs,
                     (const int *)jni::GetDirectBufferAddress(env, parameters_to_fit_buffer), estimator_id,
                     (size_t)(user_info_size < 0 ? 0 : user_info_size),
                     jni::GetDirectBufferAddress(env, user_info_buffer),
                     (float *)jni::GetDirectBufferAddress(env, output_parameters_buffer),
                     (int *)jni::GetDirectBufferAddress(env, output_states_buffer),
                     (float *)jni::GetDirectBufferAddress(env, output_chi_squares_buffer),
                     (int *)jni::GetDirectBufferAddress(env, output_number_iterations_buffer));
}

JNIEXPORT jstring JNICALL Java_gpufitImFCS_GpufitImFCS_getLastError(JNIEnv *env, jclass) {
    return jni::NewStringUTF(env, imfcs_get_last_error());
}

JNIEXPORT jboolean JNICALL Java_gpufitImFCS_GpufitImFCS_isCudaAvailableInt(JNIEnv *, jclass) {
    return imfcs_device_available() == 1 ? JNI_TRUE : JNI_FALSE;
}

JNIEXPORT jintArray JNICALL Java_gpufitImFCS_GpufitImFCS_getCudaVersionAsArray(JNIEnv *env, jclass) {
    int rv = 0, dv = 0;
    if (imfcs_get_versions(&rv, &dv) != 0) return nullptr;
    jintArray a = jni::NewIntArray(env, 2);
    jint fill[2] = {rv, dv};
    jni::SetIntArrayRegion(env, a, 0, 2, fill);
    return a;
}

JNIEXPORT void JNICALL Java_gpufitImFCS_GpufitImFCS_resetGPU(JNIEnv *env, jclass) {
    if (imfcs_reset_gpu() != 0) throw_java(env, imfcs_get_last_error());
}

JNIEXPORT void JNICALL Java_gpufitImFCS_GpufitImFCS_resetDevice(JNIEnv *, jclass) { imfcs_reset_device(); }

JNIEXPORT void JNICALL Java_gpufitImFCS_GpufitImFCS_calcDataBleachCorrection(JNIEnv *env, jclass, jfloatArray pixels,
                                                                             jfloatArray outdata, jobject params) {
    imfcs_acf_params p;
    read_params(env, params, &p);
    jfloat *in = jni::GetFloatArrayElements(env, pixels);
    size_t nout = (size_t)p.w_temp * p.h_temp * (p.nopit > 0 ? p.nopit : 0);
    float *out = new float[nout > 0 ? nout : 1];
    int rc = imfcs_calc_data_bleach_correction(in, out, &p);
    if (rc == 0) jni::SetFloatArrayRegion(env, outdata, 0, (jsize)nout, out);
    delete[] out;
    jni::ReleaseFloatArrayElements(env, pixels, in, 0);
    if (rc != 0) throw_java(env, imfcs_get_last_error());
}

JNIEXPORT jboolean JNICALL Java_gpufitImFCS_GpufitImFCS_isBinningMemorySufficient(JNIEnv *env, jclass,
                                                                                  jobject params) {
    imfcs_acf_params p;
    read_params(env, params, &p);
    int rc = imfcs_is_binning_memory_sufficient(&p);
    if (rc < 0) {
        throw_java(env, imfcs_get_last_error());
        return JNI_FALSE;
    }
    return rc == 1 ? JNI_TRUE : JNI_FALSE;
}

JNIEXPORT void JNICALL Java_gpufitImFCS_GpufitImFCS_calcBinning(JNIEnv *env, jclass, jfloatArray indata,
                                                                jfloatArray outdata, jobject params) {
    imfcs_acf_params p;
    read_params(env, params, &p);
    jfloat *in = jni::GetFloatArrayElements(env, indata);
    size_t T = (size_t)(p.lastframe - p.firstframe + 1);
    size_t nout = (size_t)p.w_temp * p.h_temp * T;
    float *out = new float[nout > 0 ? nout : 1];
    int rc = imfcs_calc_binning(in, out, &p);
    if (rc == 0) jni::SetFloatArrayRegion(env, outdata, 0, (jsize)nout, out);
    delete[] out;
    jni::ReleaseFloatArrayElements(env, indata, in, 0);
    if (rc != 0) throw_java(env, imfcs_get_last_error());
}

JNIEXPORT jboolean JNICALL Java_gpufitImFCS_GpufitImFCS_isACFmemorySufficient(JNIEnv *env, jclass, jobject params) {
    imfcs_acf_params p;
    read_params(env, params, &p);
    int rc = imfcs_is_acf_memory_sufficient(&p);
    if (rc < 0) {
        throw_java(env, imfcs_get_last_error());
        return JNI_FALSE;
    }
    return rc == 1 ? JNI_TRUE : JNI_FALSE;
}

JNIEXPORT void JNICALL Java_gpufitImFCS_GpufitImFCS_calcACF(JNIEnv *env, jclass, jfloatArray pixels,
                                                            jdoubleArray pixels1, jdoubleArray blockvararray,
                                                            jdoubleArray NBmeanGPU, jdoubleArray NBcovarianceGPU,
                                                            jdoubleArray blocked1D, jdoubleArray bleachcorr_params,
                                                            jdoubleArray samp, jintArray lag, jobject params) {
    imfcs_acf_params p;
    read_params(env, params, &p);
    size_t np = (size_t)p.width * p.height, nc = (size_t)(p.chanum > 0 ? p.chanum : 0);
    jfloat *in = jni::GetFloatArrayElements(env, pixels);
    jdouble *bc = jni::GetDoubleArrayElements(env, bleachcorr_params);
    jdouble *sp = jni::GetDoubleArrayElements(env, samp);
    jint *lg = jni::GetIntArrayElements(env, lag);
    double *o1 = new double[np * nc + 1], *o2 = new double[np * nc + 1], *o3 = new double[np * nc + 1];
    double *nbm = new double[np + 1], *nbc = new double[np + 1];
    int rc = imfcs_calc_acf(in, o1, o2, nbm, nbc, o3, bc, sp, lg, &p);
    if (rc == 0) {
        jni::SetDoubleArrayRegion(env, pixels1, 0, (jsize)(np * nc), o1);
        jni::SetDoubleArrayRegion(env, blockvararray, 0, (jsize)(np * nc), o2);
        jni::SetDoubleArrayRegion(env, blocked1D, 0, (jsize)(np * nc), o3);
        if (p.isNBcalculation) {
            jni::SetDoubleArrayRegion(env, NBmeanGPU, 0, (jsize)np, nbm);
            jni::SetDoubleArrayRegion(env, NBcovarianceGPU, 0, (jsize)np, nbc);
        }
    }
    delete[] o1;
    delete[] o2;
    delete[] o3;
    delete[] nbm;
    delete[] nbc;
    jni::ReleaseFloatArrayElements(env, pixels, in, 0);
    jni::ReleaseDoubleArrayElements(env, bleachcorr_params, bc, 0);
    jni::ReleaseDoubleArrayElements(env, samp, sp, 0);
    jni::ReleaseIntArrayElements(env, lag, lg, 0);
    if (rc != 0) throw_java(env, imfcs_get_last_error());
}

}  // extern "C"
