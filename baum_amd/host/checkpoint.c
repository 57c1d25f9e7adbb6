/* checkpoint.c -- checkpoint and restart of device-resident levels in the
 * reference's file format.
 *
 * checkpoint_write_local (src/utility/checkpoint/checkpoint_write.c:91-205):
 *   "$BEGIN_checkpoint_header:  $lmin = L , $lmax = M , $time = T\n" (time of
 *   level lmin, %.3f), one "level l : iteration = i , time = t\n" line per
 *   level (%g), "$END_checkpoint_header\n\n"; the parameter database
 *   "$BEGIN_parameter_database:  $number_of_parameters = n\n", "name = value\n"
 *   lines, "$END_parameter_database\n\n"; then "$BEGIN_variables:\n" and per
 *   level and enabled variable "$variable = name : level = l , level->nnodes
 *   = n\n", the n doubles in binary and "\n"; "$END_variables\n\n".
 * checkpoint_read_ParsAndIterations_local / checkpoint_read_Vars_local
 * (checkpoint_read.c:96-275): iterations and times per level; variables by
 * name, a point count that differs from the level's is an error, variables
 * the caller does not list (or has not enabled) are skipped.  The parameter
 * database is written as given and skipped on reading: parameters belong to
 * the caller.  Device data go through host copies of one level array at a
 * time.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/bamhip.h"

static int write_levels(FILE *fp, const bamhip_checkpoint *C)
{
  fprintf(fp, "$BEGIN_checkpoint_header:  ");
  fprintf(fp, "$lmin = %d , $lmax = %d , ", C->lmin, C->lmax);
  fprintf(fp, "$time = %.3f\n", C->time[0]);
  for (int l = C->lmin; l <= C->lmax; l++)
    fprintf(fp, "level %d : iteration = %d , time = %g\n", l, C->iteration[l - C->lmin], C->time[l - C->lmin]);
  fprintf(fp, "$END_checkpoint_header\n\n");
  fprintf(fp, "$BEGIN_parameter_database:  ");
  fprintf(fp, "$number_of_parameters = %d\n", C->nparams);
  for (int i = 0; i < C->nparams; i++) {
    fprintf(fp, "%s = ", C->param_names[i]);
    fprintf(fp, "%s\n", C->param_values[i]);
  }
  fprintf(fp, "$END_parameter_database\n\n");
  fprintf(fp, "$BEGIN_variables:\n");
  for (int l = C->lmin; l <= C->lmax; l++) {
    bamhip_level *lev = C->levels[l - C->lmin];
    long nn = 0;
    int e;
    if ((e = bamhip_level_npoints(lev, &nn))) return e;
    double *buf = malloc(sizeof(double) * (nn ? nn : 1));
    if (!buf) return BAMHIP_EINVAL;
    for (int k = 0; k < C->nvars; k++) {
      if (!bamhip_level_ptr(lev, C->vars[k])) continue; /* no storage on this level */
      if ((e = bamhip_level_get(lev, C->vars[k], buf))) {
        free(buf);
        return e;
      }
      fprintf(fp, "$variable = %s : level = %d , level->nnodes = %ld\n", C->var_names[k], l, nn);
      fwrite(buf, sizeof(double), (size_t)nn, fp);
      fprintf(fp, "\n");
    }
    free(buf);
  }
  fprintf(fp, "$END_variables\n\n");
  return BAMHIP_OK;
}

int bamhip_checkpoint_write(const char *path, const bamhip_checkpoint *C)
{
  if (!path || !C || !C->levels || !C->iteration || !C->time || C->lmax < C->lmin || C->nvars < 0 ||
      (C->nvars && (!C->vars || !C->var_names)) || (C->nparams && (!C->param_names || !C->param_values)))
    return BAMHIP_EINVAL;
  FILE *fp = fopen(path, "wb");
  if (!fp) return BAMHIP_EINVAL;
  const int e = write_levels(fp, C);
  if (fclose(fp) && !e) return BAMHIP_EINVAL;
  return e;
}

/* the text after "key = " on a line, up to the next " ," or the end */
static int value_of(const char *line, const char *key, char *out, size_t n)
{
  const char *p = strstr(line, key);
  if (!p) return 0;
  p += strlen(key);
  while (*p == ' ') p++;
  if (*p != '=') return 0;
  p++;
  while (*p == ' ') p++;
  size_t i = 0;
  while (p[i] && p[i] != '\n' && !(p[i] == ' ' && p[i + 1] == ',') && !(p[i] == ' ' && p[i + 1] == ':') &&
         i + 1 < n) {
    out[i] = p[i];
    i++;
  }
  out[i] = 0;
  return 1;
}

int bamhip_checkpoint_read(const char *path, const bamhip_checkpoint *C, int *iteration, double *time)
{
  char line[4096], val[1024];
  int e = BAMHIP_OK;
  if (!path || !C || !C->levels || C->lmax < C->lmin) return BAMHIP_EINVAL;
  FILE *fp = fopen(path, "rb");
  if (!fp) return BAMHIP_EINVAL;
  double *buf = 0;
  /* header: iterations and times per level */
  while (fgets(line, sizeof line, fp)) {
    if (!strncmp(line, "$END_checkpoint_header", 22)) break;
    if (!strncmp(line, "level ", 6)) {
      const int l = atoi(line + 6);
      if (l >= C->lmin && l <= C->lmax) {
        if (iteration && value_of(line, "iteration", val, sizeof val)) iteration[l - C->lmin] = atoi(val);
        if (time && value_of(line, "time", val, sizeof val)) time[l - C->lmin] = atof(val);
      }
    }
  }
  /* the header's time is for human inspection (%g); the reference restarts
   * with time = iteration * dt (checkpoint_read.c:117-124) */
  if (time && iteration && C->dt)
    for (int l = C->lmin; l <= C->lmax; l++) time[l - C->lmin] = iteration[l - C->lmin] * C->dt[l - C->lmin];
  /* variables */
  while (fgets(line, sizeof line, fp))
    if (!strncmp(line, "$BEGIN_variables:", 17)) break;
  while (fgets(line, sizeof line, fp)) {
    if (!strncmp(line, "$END_variables", 14)) break;
    if (strncmp(line, "$variable", 9)) continue;
    char name[1024], lv[64], nn[64];
    if (!value_of(line, "$variable", name, sizeof name) || !value_of(line, "level", lv, sizeof lv) ||
        !value_of(line, "level->nnodes", nn, sizeof nn)) {
      e = BAMHIP_EINVAL;
      break;
    }
    const int l = atoi(lv);
    const long n = atol(nn);
    int var = -1;
    for (int k = 0; k < C->nvars; k++)
      if (!strcmp(C->var_names[k], name)) var = C->vars[k];
    bamhip_level *lev = (l >= C->lmin && l <= C->lmax) ? C->levels[l - C->lmin] : 0;
    if (lev && var >= 0 && bamhip_level_ptr(lev, var)) {
      long np = 0;
      bamhip_level_npoints(lev, &np);
      if (np != n) { /* level->nnodes in checkpoint file is not what it should be */
        e = BAMHIP_EINVAL;
        break;
      }
      free(buf);
      buf = malloc(sizeof(double) * (n ? n : 1));
      if (!buf || fread(buf, sizeof(double), (size_t)n, fp) != (size_t)n) {
        e = BAMHIP_EINVAL;
        break;
      }
      if ((e = bamhip_level_put(lev, var, buf))) break;
    } else if (fseek(fp, (long)(n * (long)sizeof(double)), SEEK_CUR)) { /* skipped */
      e = BAMHIP_EINVAL;
      break;
    }
    if (fgetc(fp) != '\n') { /* the newline after the block */
      e = BAMHIP_EINVAL;
      break;
    }
  }
  free(buf);
  fclose(fp);
  return e;
}
