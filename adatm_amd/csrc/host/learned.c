/*
 * Learned memoization planner.
 *
 * Beside the reference's analytic model (min predicted products,
 * src/adatm_bench.c:108-142) a small MLP predicts the GPU time of the
 * MTTKRP sequence of each candidate plan from 24 statistics of the tensor
 * and the plan, and the candidate with the smallest prediction is chosen.
 * The network (24 -> 64 -> 64 -> 64 -> 1, FP32, ReLU) is trained on the GPU
 * by the HIP kernels of mlp.hip on timings the engine itself produces on
 * synthetic tensors; without a model the analytic choice is used.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "adatm_host.h"
#include "../hip/adatm_kern.h"

#define NF ADATM_MLP_IN
#define NH ADATM_MLP_HID
#define MAXN ADATM_MAX_NMODES

/* parameter layout: W1 b1 W2 b2 W3 b3 W4 b4 */
static uint32_t const LIN[4] = {NF, NH, NH, NH};
static uint32_t const LOUT[4] = {NH, NH, NH, 1};

size_t adatm_mlp_nparams(void)
{
  size_t n = 0;
  for (int l = 0; l < 4; ++l) n += (size_t)LIN[l] * LOUT[l] + LOUT[l];
  return n;
}

static size_t woff(int l)
{
  size_t n = 0;
  for (int k = 0; k < l; ++k) n += (size_t)LIN[k] * LOUT[k] + LOUT[k];
  return n;
}

struct adatm_mlp {
  int device;
  hipStream_t s;
  float *p, *g, *m, *v; /* device, nparams each */
  float xmean[NF], xstd[NF], ymean, ystd;
  uint32_t adam_t;
};

static int merr(char const *what)
{
  adatm_set_error("mlp: %s failed", what);
  return SPLATT_ERROR_DEVICE;
}

/* splitmix64 stream for initialisation and shuffles */
static uint64_t sm64(uint64_t *st)
{
  uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double sm_unif(uint64_t *st) { return (sm64(st) >> 11) * (1.0 / 9007199254740992.0); }

int adatm_mlp_create(int device, uint64_t seed, adatm_mlp **out)
{
  *out = NULL;
  adatm_mlp *M = (adatm_mlp *)calloc(1, sizeof(adatm_mlp));
  if (!M) return SPLATT_ERROR_NOMEMORY;
  M->device = device;
  size_t const np = adatm_mlp_nparams();
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&M->s, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc((void **)&M->p, np * 4) != hipSuccess ||
      hipMalloc((void **)&M->g, np * 4) != hipSuccess ||
      hipMalloc((void **)&M->m, np * 4) != hipSuccess ||
      hipMalloc((void **)&M->v, np * 4) != hipSuccess) {
    adatm_mlp_destroy(M);
    return merr("allocation");
  }
  /* He-uniform weights, zero biases */
  float *h = (float *)calloc(np, sizeof(float));
  uint64_t st = seed ^ 0xA5A5A5A5ull;
  for (int l = 0; l < 4; ++l) {
    float const a = sqrtf(6.0f / (float)LIN[l]);
    float *w = h + woff(l);
    for (size_t i = 0; i < (size_t)LIN[l] * LOUT[l]; ++i)
      w[i] = (float)((2.0 * sm_unif(&st) - 1.0) * a);
  }
  int rc = adatm_mlp_set_params(M, h, np);
  free(h);
  for (int i = 0; i < NF; ++i) {
    M->xmean[i] = 0.f;
    M->xstd[i] = 1.f;
  }
  M->ymean = 0.f;
  M->ystd = 1.f;
  if (rc != SPLATT_SUCCESS) {
    adatm_mlp_destroy(M);
    return rc;
  }
  *out = M;
  return SPLATT_SUCCESS;
}

void adatm_mlp_destroy(adatm_mlp *M)
{
  if (!M) return;
  hipFree(M->p);
  hipFree(M->g);
  hipFree(M->m);
  hipFree(M->v);
  if (M->s) hipStreamDestroy(M->s);
  free(M);
}

int adatm_mlp_set_params(adatm_mlp *M, float const *in, size_t n)
{
  if (!M || n != adatm_mlp_nparams()) return SPLATT_ERROR_BADINPUT;
  if (hipMemcpy(M->p, in, n * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(M->m, 0, n * 4) != hipSuccess ||
      hipMemset(M->v, 0, n * 4) != hipSuccess)
    return merr("set params");
  M->adam_t = 0;
  return SPLATT_SUCCESS;
}

int adatm_mlp_get_params(adatm_mlp *M, float *out, size_t n)
{
  if (!M || n != adatm_mlp_nparams()) return SPLATT_ERROR_BADINPUT;
  if (hipStreamSynchronize(M->s) != hipSuccess ||
      hipMemcpy(out, M->p, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return merr("get params");
  return SPLATT_SUCCESS;
}

int adatm_mlp_set_norm(adatm_mlp *M, float const *xmean, float const *xstd,
                       float ymean, float ystd)
{
  if (!M) return SPLATT_ERROR_BADINPUT;
  memcpy(M->xmean, xmean, sizeof(M->xmean));
  memcpy(M->xstd, xstd, sizeof(M->xstd));
  M->ymean = ymean;
  M->ystd = ystd;
  return SPLATT_SUCCESS;
}

int adatm_mlp_get_norm(adatm_mlp *M, float *xmean, float *xstd, float *ymean,
                       float *ystd)
{
  if (!M) return SPLATT_ERROR_BADINPUT;
  memcpy(xmean, M->xmean, sizeof(M->xmean));
  memcpy(xstd, M->xstd, sizeof(M->xstd));
  *ymean = M->ymean;
  *ystd = M->ystd;
  return SPLATT_SUCCESS;
}

/* activations of one batch on the device */
typedef struct {
  uint32_t B;
  float *x, *t;         /* B x NF standardized inputs, B targets */
  float *Z[4], *H[4];   /* pre-activations / activations per layer */
  float *dA, *dB2;      /* B x NH gradient buffers */
  float *part;
} mlp_batch;

static void batch_free(mlp_batch *b)
{
  hipFree(b->x);
  hipFree(b->t);
  for (int l = 0; l < 4; ++l) {
    hipFree(b->Z[l]);
    hipFree(b->H[l]);
  }
  hipFree(b->dA);
  hipFree(b->dB2);
  hipFree(b->part);
  memset(b, 0, sizeof(*b));
}

static int batch_alloc(mlp_batch *b, uint32_t B)
{
  memset(b, 0, sizeof(*b));
  b->B = B;
  int bad = hipMalloc((void **)&b->x, (size_t)B * NF * 4) != hipSuccess;
  bad |= hipMalloc((void **)&b->t, (size_t)B * 4) != hipSuccess;
  for (int l = 0; l < 4; ++l) {
    bad |= hipMalloc((void **)&b->Z[l], (size_t)B * LOUT[l] * 4) != hipSuccess;
    bad |= hipMalloc((void **)&b->H[l], (size_t)B * LOUT[l] * 4) != hipSuccess;
  }
  bad |= hipMalloc((void **)&b->dA, (size_t)B * NH * 4) != hipSuccess;
  bad |= hipMalloc((void **)&b->dB2, (size_t)B * NH * 4) != hipSuccess;
  bad |= hipMalloc((void **)&b->part, ((size_t)B / 256 + 2) * 4) != hipSuccess;
  if (bad) {
    batch_free(b);
    return SPLATT_ERROR_NOMEMORY;
  }
  return SPLATT_SUCCESS;
}

/* forward pass of B standardised rows x (device, B x NF) on the network's
 * stream: Z[l] / H[l] (device, B x LOUT[l]) receive every layer's
 * pre-activations and activations; the prediction is H[3] */
int adatm_mlp_forward_dev(adatm_mlp *M, float const *x, float *const *Z,
                          float *const *H, uint32_t B)
{
  float const *in = x;
  for (int l = 0; l < 4; ++l) {
    size_t const o = woff(l);
    if (adatm_k_dense_fwd(M->s, in, M->p + o, M->p + o + (size_t)LIN[l] * LOUT[l],
                          Z[l], H[l], B, LIN[l], LOUT[l], l < 3))
      return merr("forward");
    in = H[l];
  }
  return SPLATT_SUCCESS;
}

/* the network's device, stream and standardisation */
void adatm_mlp_dev_info(adatm_mlp const *M, int *device, hipStream_t *stream,
                        float *xmean, float *xstd, float *ymean, float *ystd)
{
  *device = M->device;
  *stream = M->s;
  memcpy(xmean, M->xmean, sizeof(M->xmean));
  memcpy(xstd, M->xstd, sizeof(M->xstd));
  *ymean = M->ymean;
  *ystd = M->ystd;
}

static int forward(adatm_mlp *M, mlp_batch *b, uint32_t B)
{
  return adatm_mlp_forward_dev(M, b->x, b->Z, b->H, B);
}

/* gradient of the batch MSE into M->g; *loss (nullable) = mean squared error */
static int backward(adatm_mlp *M, mlp_batch *b, uint32_t B, float *loss)
{
  /* dy in dB2 (B x 1) */
  if (adatm_k_mse(M->s, b->H[3], b->t, B, b->dB2, b->part)) return merr("mse");
  float *dz = b->dB2, *dh = b->dA;
  for (int l = 3; l >= 0; --l) {
    size_t const o = woff(l);
    float const *Hin = (l == 0) ? b->x : b->H[l - 1];
    float const *Zprev = (l == 0) ? NULL : b->Z[l - 1];
    if (adatm_k_dense_bwd(M->s, Hin, dz, M->p + o, Zprev, l > 0 ? dh : NULL,
                          M->g + o, M->g + o + (size_t)LIN[l] * LOUT[l], B, LIN[l],
                          LOUT[l]))
      return merr("backward");
    float *t = dz;
    dz = dh;
    dh = t;
  }
  if (loss) {
    uint32_t const nb = (B + 255) / 256;
    float part[4096];
    if (nb > 4096 || hipMemcpyAsync(part, b->part, nb * 4, hipMemcpyDeviceToHost,
                                    M->s) != hipSuccess ||
        hipStreamSynchronize(M->s) != hipSuccess)
      return merr("loss");
    double s = 0.0;
    for (uint32_t i = 0; i < nb; ++i) s += part[i];
    *loss = (float)(s / (double)B);
  }
  return SPLATT_SUCCESS;
}

static void standardize(adatm_mlp const *M, float const *X, uint32_t n, float *out)
{
  for (uint32_t r = 0; r < n; ++r)
    for (int i = 0; i < NF; ++i)
      out[(size_t)r * NF + i] = (X[(size_t)r * NF + i] - M->xmean[i]) / M->xstd[i];
}

int adatm_mlp_loss_grad(adatm_mlp *M, float const *X, float const *y,
                        uint32_t n, float *grad, float *loss)
{
  if (!M || n == 0) return SPLATT_ERROR_BADINPUT;
  mlp_batch b;
  if (batch_alloc(&b, n) != SPLATT_SUCCESS) return SPLATT_ERROR_NOMEMORY;
  float *xs = (float *)malloc((size_t)n * NF * 4);
  float *ts = (float *)malloc((size_t)n * 4);
  standardize(M, X, n, xs);
  for (uint32_t r = 0; r < n; ++r) ts[r] = (y[r] - M->ymean) / M->ystd;
  int rc = SPLATT_SUCCESS;
  if (hipMemcpy(b.x, xs, (size_t)n * NF * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b.t, ts, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = merr("upload");
  if (rc == SPLATT_SUCCESS) rc = forward(M, &b, n);
  if (rc == SPLATT_SUCCESS) rc = backward(M, &b, n, loss);
  if (rc == SPLATT_SUCCESS &&
      (hipStreamSynchronize(M->s) != hipSuccess ||
       hipMemcpy(grad, M->g, adatm_mlp_nparams() * 4, hipMemcpyDeviceToHost) !=
           hipSuccess))
    rc = merr("grad download");
  free(xs);
  free(ts);
  batch_free(&b);
  return rc;
}

/* one Adam step on (X, y) with the current standardisation */
int adatm_mlp_step(adatm_mlp *M, float const *X, float const *y, uint32_t n,
                   float lr, float *loss)
{
  if (!M || n == 0) return SPLATT_ERROR_BADINPUT;
  mlp_batch b;
  if (batch_alloc(&b, n) != SPLATT_SUCCESS) return SPLATT_ERROR_NOMEMORY;
  float *xs = (float *)malloc((size_t)n * NF * 4);
  float *ts = (float *)malloc((size_t)n * 4);
  standardize(M, X, n, xs);
  for (uint32_t r = 0; r < n; ++r) ts[r] = (y[r] - M->ymean) / M->ystd;
  int rc = SPLATT_SUCCESS;
  if (hipMemcpy(b.x, xs, (size_t)n * NF * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b.t, ts, (size_t)n * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = merr("upload");
  if (rc == SPLATT_SUCCESS) rc = forward(M, &b, n);
  if (rc == SPLATT_SUCCESS) rc = backward(M, &b, n, loss);
  if (rc == SPLATT_SUCCESS) {
    float const b1 = 0.9f, b2 = 0.999f;
    M->adam_t++;
    float const c1 = 1.f - powf(b1, (float)M->adam_t);
    float const c2 = 1.f - powf(b2, (float)M->adam_t);
    if (adatm_k_adam(M->s, M->p, M->g, M->m, M->v, (uint32_t)adatm_mlp_nparams(), lr,
                     b1, b2, 1e-8f, c1, c2) ||
        hipStreamSynchronize(M->s) != hipSuccess)
      rc = merr("adam");
  }
  free(xs);
  free(ts);
  batch_free(&b);
  return rc;
}

int adatm_mlp_train(adatm_mlp *M, float const *X, float const *y, uint32_t n,
                    uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                    float *loss_hist)
{
  if (!M || n == 0 || batch == 0) return SPLATT_ERROR_BADINPUT;
  /* standardisation from the training set */
  for (int i = 0; i < NF; ++i) {
    double s = 0.0, s2 = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
      double const v = X[(size_t)r * NF + i];
      s += v;
      s2 += v * v;
    }
    double const mu = s / n, var = s2 / n - mu * mu;
    M->xmean[i] = (float)mu;
    M->xstd[i] = (float)(var > 1e-12 ? sqrt(var) : 1.0);
  }
  {
    double s = 0.0, s2 = 0.0;
    for (uint32_t r = 0; r < n; ++r) {
      s += y[r];
      s2 += (double)y[r] * y[r];
    }
    double const mu = s / n, var = s2 / n - mu * mu;
    M->ymean = (float)mu;
    M->ystd = (float)(var > 1e-12 ? sqrt(var) : 1.0);
  }
  uint32_t const B = batch < n ? batch : n;
  mlp_batch b;
  if (batch_alloc(&b, B) != SPLATT_SUCCESS) return SPLATT_ERROR_NOMEMORY;
  float *xs = (float *)malloc((size_t)n * NF * 4);
  float *ts = (float *)malloc((size_t)n * 4);
  float *xb = (float *)malloc((size_t)B * NF * 4);
  float *tb = (float *)malloc((size_t)B * 4);
  uint32_t *perm = (uint32_t *)malloc((size_t)n * 4);
  standardize(M, X, n, xs);
  for (uint32_t r = 0; r < n; ++r) {
    ts[r] = (y[r] - M->ymean) / M->ystd;
    perm[r] = r;
  }
  uint64_t st = seed;
  size_t const np = adatm_mlp_nparams();
  int rc = SPLATT_SUCCESS;
  float const b1 = 0.9f, b2 = 0.999f;
  for (uint32_t ep = 0; ep < epochs && rc == SPLATT_SUCCESS; ++ep) {
    for (uint32_t i = n - 1; i > 0; --i) { /* Fisher-Yates */
      uint32_t const j = (uint32_t)(sm64(&st) % (i + 1));
      uint32_t const t = perm[i];
      perm[i] = perm[j];
      perm[j] = t;
    }
    double lsum = 0.0;
    uint32_t nbat = 0;
    for (uint32_t s0 = 0; s0 + B <= n && rc == SPLATT_SUCCESS; s0 += B) {
      for (uint32_t r = 0; r < B; ++r) {
        memcpy(xb + (size_t)r * NF, xs + (size_t)perm[s0 + r] * NF, NF * 4);
        tb[r] = ts[perm[s0 + r]];
      }
      if (hipMemcpyAsync(b.x, xb, (size_t)B * NF * 4, hipMemcpyHostToDevice, M->s) !=
              hipSuccess ||
          hipMemcpyAsync(b.t, tb, (size_t)B * 4, hipMemcpyHostToDevice, M->s) !=
              hipSuccess) {
        rc = merr("batch upload");
        break;
      }
      float l = 0.f;
      rc = forward(M, &b, B);
      if (rc == SPLATT_SUCCESS) rc = backward(M, &b, B, &l);
      if (rc != SPLATT_SUCCESS) break;
      M->adam_t++;
      float const c1 = 1.f - powf(b1, (float)M->adam_t);
      float const c2 = 1.f - powf(b2, (float)M->adam_t);
      if (adatm_k_adam(M->s, M->p, M->g, M->m, M->v, (uint32_t)np, lr, b1, b2,
                       1e-8f, c1, c2))
        rc = merr("adam");
      lsum += l;
      ++nbat;
    }
    if (loss_hist) loss_hist[ep] = nbat ? (float)(lsum / nbat) : 0.f;
  }
  if (hipStreamSynchronize(M->s) != hipSuccess && rc == SPLATT_SUCCESS)
    rc = merr("train sync");
  free(xs);
  free(ts);
  free(xb);
  free(tb);
  free(perm);
  batch_free(&b);
  return rc;
}

/* predicted value of the target (log ms) for n raw feature rows */
int adatm_mlp_predict(adatm_mlp *M, float const *X, uint32_t n, float *out)
{
  if (!M || n == 0) return SPLATT_ERROR_BADINPUT;
  mlp_batch b;
  if (batch_alloc(&b, n) != SPLATT_SUCCESS) return SPLATT_ERROR_NOMEMORY;
  float *xs = (float *)malloc((size_t)n * NF * 4);
  standardize(M, X, n, xs);
  int rc = SPLATT_SUCCESS;
  if (hipMemcpy(b.x, xs, (size_t)n * NF * 4, hipMemcpyHostToDevice) != hipSuccess)
    rc = merr("upload");
  if (rc == SPLATT_SUCCESS) rc = forward(M, &b, n);
  if (rc == SPLATT_SUCCESS &&
      (hipStreamSynchronize(M->s) != hipSuccess ||
       hipMemcpy(out, b.H[3], (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess))
    rc = merr("predict download");
  if (rc == SPLATT_SUCCESS)
    for (uint32_t r = 0; r < n; ++r) out[r] = out[r] * M->ystd + M->ymean;
  free(xs);
  batch_free(&b);
  return rc;
}

static char const MAGIC[8] = {'A', 'D', 'M', 'L', 'P', '1', 0, 0};

int adatm_mlp_save(adatm_mlp *M, char const *path)
{
  size_t const np = adatm_mlp_nparams();
  float *p = (float *)malloc(np * 4);
  int rc = adatm_mlp_get_params(M, p, np);
  if (rc != SPLATT_SUCCESS) {
    free(p);
    return rc;
  }
  FILE *f = fopen(path, "wb");
  if (!f) {
    free(p);
    adatm_set_error("mlp: cannot write %s", path);
    return SPLATT_ERROR_BADINPUT;
  }
  uint64_t const n64 = np;
  fwrite(MAGIC, 1, 8, f);
  fwrite(&n64, 8, 1, f);
  fwrite(M->xmean, 4, NF, f);
  fwrite(M->xstd, 4, NF, f);
  fwrite(&M->ymean, 4, 1, f);
  fwrite(&M->ystd, 4, 1, f);
  fwrite(p, 4, np, f);
  fclose(f);
  free(p);
  return SPLATT_SUCCESS;
}

int adatm_mlp_load(int device, char const *path, adatm_mlp **out)
{
  *out = NULL;
  FILE *f = fopen(path, "rb");
  if (!f) {
    adatm_set_error("mlp: cannot read %s", path);
    return SPLATT_ERROR_BADINPUT;
  }
  char mg[8];
  uint64_t n64 = 0;
  size_t const np = adatm_mlp_nparams();
  float xm[NF], xsd[NF], ym, ysd;
  float *p = (float *)malloc(np * 4);
  int ok = fread(mg, 1, 8, f) == 8 && !memcmp(mg, MAGIC, 8) &&
           fread(&n64, 8, 1, f) == 1 && n64 == np &&
           fread(xm, 4, NF, f) == NF && fread(xsd, 4, NF, f) == NF &&
           fread(&ym, 4, 1, f) == 1 && fread(&ysd, 4, 1, f) == 1 &&
           fread(p, 4, np, f) == np;
  fclose(f);
  if (!ok) {
    free(p);
    adatm_set_error("mlp: %s is not a model file", path);
    return SPLATT_ERROR_BADINPUT;
  }
  adatm_mlp *M = NULL;
  int rc = adatm_mlp_create(device, 1, &M);
  if (rc == SPLATT_SUCCESS) rc = adatm_mlp_set_params(M, p, np);
  if (rc == SPLATT_SUCCESS) rc = adatm_mlp_set_norm(M, xm, xsd, ym, ysd);
  free(p);
  if (rc != SPLATT_SUCCESS) {
    adatm_mlp_destroy(M);
    return rc;
  }
  *out = M;
  return SPLATT_SUCCESS;
}

/* ---- plan features ------------------------------------------------------- */
static int memo_lvl(adatm_group const *G, uint32_t l)
{
  return G->n_imten > 0 && l + G->n_imten >= G->begin_imten && l + 1 <= G->begin_imten;
}

static uint32_t top_of(adatm_group const *G, uint32_t lb)
{
  uint32_t lt = lb - 1;
  while (lt > 0 && !memo_lvl(G, lt) && lb - lt < ADATM_KMAXFUSE) --lt;
  return lt;
}

/* bytes of a fused pass lb -> lt (FP32 path), same model as the engine */
static double pass_bytes(uint32_t N, uint64_t const *F, uint64_t const *dims,
                         uint32_t const *order, uint32_t lb, uint32_t lt, double R)
{
  double b = (lb == N - 1) ? (double)F[lb] * 8.0 : (double)F[lb] * (R * 4.0 + 4.0);
  for (uint32_t l = lt + 1; l < lb; ++l) b += (double)F[l] * 8.0 + 4.0;
  b += ((double)F[lt] + 1.0) * 4.0;
  for (uint32_t l = lt + 1; l <= lb; ++l) b += (double)dims[order[l]] * R * 4.0;
  b += (double)F[lt] * R * 4.0;
  return b;
}

/* F of every group of plan p: Fg[g*N + l] */
int adatm_plan_features_from_counts(uint32_t N, uint64_t nnz, uint64_t const *dims,
                                    uint64_t R, adatm_plan const *p,
                                    uint64_t const *Fg, float *f)
{
  double const lR = (double)R;
  uint32_t ug[MAXN], ut[MAXN];
  adatm_plan_use(p, N, ug, ut);
  double ld[MAXN], mean = 0.0, mx = -1e9, mn = 1e9;
  for (uint32_t m = 0; m < N; ++m) {
    ld[m] = log10((double)dims[m]);
    mean += ld[m];
    mx = ld[m] > mx ? ld[m] : mx;
    mn = ld[m] < mn ? ld[m] : mn;
  }
  mean /= N;
  double var = 0.0;
  for (uint32_t m = 0; m < N; ++m) var += (ld[m] - mean) * (ld[m] - mean);
  double bytes = 0.0, flops = 0.0, maxpass = 0.0, memo_rows = 0.0;
  double scat_nodes = 0.0, prop_rows = 0.0, depth_sum = 0.0;
  double leaf_len = 0.0, root_cnt = 0.0, branch = 0.0;
  uint32_t nreuse = 0, nmemo = 0, maxK = 0, launches = 0;
  for (uint32_t g = 0; g < p->n_grp; ++g) {
    adatm_group const *G = p->grp + g;
    uint64_t const *F = Fg + (size_t)g * N;
    nmemo += G->n_imten;
    for (uint32_t l = 1; l + 1 < N; ++l)
      if (memo_lvl(G, l)) memo_rows += (double)F[l];
    leaf_len += log10((double)nnz / (double)F[N - 2]);
    root_cnt += log10((double)F[0]);
    branch += log10((double)F[1] / (double)F[0]);
  }
  for (uint32_t m = 0; m < N; ++m) {
    adatm_group const *G = p->grp + ug[m];
    uint64_t const *F = Fg + (size_t)ug[m] * N;
    uint32_t const *o = G->o_opt;
    if (ut[m] == 0) {
      for (uint32_t lb = N - 1; lb > 0;) {
        uint32_t const lt = top_of(G, lb);
        double const pb = pass_bytes(N, F, dims, o, lb, lt, lR);
        bytes += pb;
        maxpass = pb > maxpass ? pb : maxpass;
        for (uint32_t l = lt + 1; l <= lb; ++l) flops += 2.0 * lR * (double)F[l];
        if (lb - lt > maxK) maxK = lb - lt;
        ++launches;
        lb = lt;
      }
    } else {
      uint32_t d = 0;
      while (o[d] != m) ++d;
      uint32_t const bd = G->begin_imten - d;
      uint32_t lb = G->begin_imten - (G->n_imten < bd ? G->n_imten : bd);
      while (lb > d) {
        uint32_t const lt = lb > d + ADATM_KMAXFUSE ? lb - ADATM_KMAXFUSE : d;
        double const pb = pass_bytes(N, F, dims, o, lb, lt, lR);
        bytes += pb;
        maxpass = pb > maxpass ? pb : maxpass;
        prop_rows += (double)F[lt];
        for (uint32_t l = lt + 1; l <= lb; ++l) flops += 2.0 * lR * (double)F[l];
        ++launches;
        lb = lt;
      }
      double sb = (double)F[d] * (lR * 4.0 + 4.0 * (d + 1)) +
                  ((double)dims[m] + 1.0) * 4.0 + (double)dims[m] * lR * 4.0;
      for (uint32_t l = 0; l < d; ++l) sb += (double)dims[o[l]] * lR * 4.0;
      bytes += sb;
      maxpass = sb > maxpass ? sb : maxpass;
      flops += lR * (double)(d + 1) * (double)F[d];
      scat_nodes += (double)F[d];
      depth_sum += d;
      ++nreuse;
      ++launches;
    }
  }
  /* the reference's analytic model on fresh statistics (no quirks) */
  uint64_t ps = 0, po = 0;
  adatm_plan_model(N, nnz, R, p, Fg, &ps, &po);
  double const pops = (double)po, pspace = (double)ps;
  double const ng = (double)p->n_grp;
  f[0] = (float)N;
  f[1] = (float)log10((double)nnz);
  f[2] = (float)log2(lR);
  f[3] = (float)mean;
  f[4] = (float)mx;
  f[5] = (float)mn;
  f[6] = (float)sqrt(var / N);
  f[7] = (float)p->n_grp;
  f[8] = (float)nmemo;
  f[9] = (float)nreuse;
  f[10] = (float)log10(bytes);
  f[11] = (float)log10(flops);
  f[12] = (float)log10(pops > 1.0 ? pops : 1.0);
  f[13] = (float)log10(pspace);
  f[14] = (float)log10(1.0 + memo_rows);
  f[15] = (float)(leaf_len / ng);
  f[16] = (float)(root_cnt / ng);
  f[17] = (float)maxK;
  f[18] = (float)log10(1.0 + scat_nodes);
  f[19] = (float)(nreuse ? depth_sum / nreuse : 0.0);
  f[20] = (float)log10(1.0 + prop_rows);
  f[21] = (float)(branch / ng);
  f[22] = (float)log10(maxpass);
  f[23] = (float)launches;
  return SPLATT_SUCCESS;
}

/* features of every candidate plan (counts on the device when device >= 0) */
int adatm_plan_features(adatm_coo const *t, uint64_t R, adatm_plan const *cands,
                        int nc, int device, float *feat)
{
  uint32_t const N = t->nmodes;
  uint32_t *orders = (uint32_t *)malloc((size_t)nc * MAXN * N * 4);
  uint32_t no = 0;
  for (int c = 0; c < nc; ++c)
    for (uint32_t g = 0; g < cands[c].n_grp; ++g) {
      int dup = 0;
      for (uint32_t k = 0; k < no && !dup; ++k)
        dup = !memcmp(orders + (size_t)k * N, cands[c].grp[g].o_opt, N * 4);
      if (!dup) memcpy(orders + (size_t)(no++) * N, cands[c].grp[g].o_opt, N * 4);
    }
  uint64_t *F = (uint64_t *)malloc((size_t)(no ? no : 1) * N * 8);
  int rc = SPLATT_SUCCESS;
  if (device >= 0)
    rc = adatm_count_fibers_device_many(t, orders, no, F, device);
  else
    for (uint32_t k = 0; rc == SPLATT_SUCCESS && k < no; ++k)
      rc = adatm_count_fibers_host(t, orders + (size_t)k * N, F + (size_t)k * N);
  uint64_t Fg[MAXN * MAXN];
  for (int c = 0; rc == SPLATT_SUCCESS && c < nc; ++c) {
    for (uint32_t g = 0; g < cands[c].n_grp; ++g)
      for (uint32_t k = 0; k < no; ++k)
        if (!memcmp(orders + (size_t)k * N, cands[c].grp[g].o_opt, N * 4)) {
          memcpy(Fg + (size_t)g * N, F + (size_t)k * N, N * 8);
          break;
        }
    rc = adatm_plan_features_from_counts(N, t->nnz, t->dims, R, cands + c, Fg,
                                         feat + (size_t)c * NF);
  }
  free(orders);
  free(F);
  return rc;
}

/* Learned choice: the candidate with the smallest predicted time; the
 * analytic choice (with the reference's statistics reuse) when model is NULL.
 * pred (nullable, one per candidate) receives the predictions (log ms). */
int adatm_plan_learned(adatm_coo const *t, uint64_t R, adatm_mlp *model,
                       int device, adatm_plan *cands, int maxc, int *ncands,
                       int *chosen, float *pred)
{
  int rc = adatm_plan_analytic(t, R, 1, (double)SPLATT_ADATM_SLIMIT, device, cands,
                               maxc, ncands, chosen);
  if (rc != SPLATT_SUCCESS || !model) return rc;
  int const nc = *ncands;
  float *feat = (float *)malloc((size_t)nc * NF * 4);
  float *y = (float *)malloc((size_t)nc * 4);
  rc = adatm_plan_features(t, R, cands, nc, device, feat);
  if (rc == SPLATT_SUCCESS) rc = adatm_mlp_predict(model, feat, (uint32_t)nc, y);
  if (rc == SPLATT_SUCCESS) {
    int best = 0;
    for (int c = 1; c < nc; ++c)
      if (y[c] < y[best]) best = c;
    *chosen = best;
    if (pred) memcpy(pred, y, (size_t)nc * 4);
  }
  free(feat);
  free(y);
  return rc;
}

/* Exhaustive search: every candidate built on the device and one CP-ALS
 * sweep (hipGraph) timed, median of nrep after nwarm; a candidate that
 * cannot be built (device memory) gets +inf.  ms (nullable) receives the
 * timings. */
int adatm_plan_exhaustive(adatm_coo const *t, uint64_t R, int precision,
                          int device, int nwarm, int nrep, adatm_plan *cands,
                          int maxc, int *ncands, int *chosen, double *ms)
{
  int ach = 0;
  int rc = adatm_plan_analytic(t, R, 1, (double)SPLATT_ADATM_SLIMIT, device, cands,
                               maxc, ncands, &ach);
  if (rc != SPLATT_SUCCESS) return rc;
  int best = -1;
  double bt = INFINITY;
  for (int c = 0; c < *ncands; ++c) {
    adatm_engine *e = NULL;
    double m = INFINITY;
    if (adatm_engine_create(t, R, cands + c, precision, device, &e) == SPLATT_SUCCESS) {
      if (adatm_engine_init_factors(e, 1) != SPLATT_SUCCESS ||
          adatm_engine_time(e, 1, nwarm, nrep, 1, &m, NULL) != SPLATT_SUCCESS)
        m = INFINITY;
      adatm_engine_destroy(e);
    }
    if (ms) ms[c] = m;
    if (m < bt) {
      bt = m;
      best = c;
    }
  }
  if (best < 0) {
    adatm_set_error("exhaustive planner: no candidate could be built on device %d",
                    device);
    return SPLATT_ERROR_NOMEMORY;
  }
  adatm_set_error("%s", "");
  *chosen = best;
  return SPLATT_SUCCESS;
}

/* The cost model file: $ADATM_COSTMODEL, else planner/costmodel.bin next to
 * the library's directory (lib/../planner).  Returns 0 when there is none. */
int adatm_costmodel_path(char *buf, size_t n)
{
  char const *env = getenv("ADATM_COSTMODEL");
  if (env && env[0]) {
    snprintf(buf, n, "%s", env);
    return 1;
  }
  Dl_info info;
  if (!dladdr((void *)&adatm_costmodel_path, &info) || !info.dli_fname) return 0;
  char dir[4096];
  snprintf(dir, sizeof(dir), "%s", info.dli_fname);
  char *slash = strrchr(dir, '/');
  if (!slash) return 0;
  *slash = 0;
  snprintf(buf, n, "%s/../planner/costmodel.bin", dir);
  FILE *f = fopen(buf, "rb");
  if (!f) return 0;
  fclose(f);
  return 1;
}
