/* File formats on either side of the E-step: STAR metadata tables and MRC images, stacks and maps.
 *
 * Host-side C ABI of librln_estep.so (no GPU work). Every call returns RLN_OK (0) or a negative RLN_ERR_* status
 * (rln_estep.h); rln_last_error() gives the message. Values of STAR tables are kept as the text the file holds and
 * converted on access, so that a table read and written again keeps its values verbatim.
 */
#ifndef RLN_FORMATS_H
#define RLN_FORMATS_H

#include <stddef.h>

#include "rln_estep.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------------
 * STAR files: MetaDataTable (src/metadata_table.h), readAll / readStar / readStarLoop / readStarList
 * (src/metadata_table.cpp:1036-1345) and write (:1370-1500); tokenizer nextTokenInSTAR (src/strings.cpp:595-662).
 * ------------------------------------------------------------------------------------------- */
typedef struct rln_star rln_star;

/* MetaDataTable::readAll: every data_ block of the file, loop_ tables and label-value lists (a list followed by a
 * loop in one block keeps the list). Errors as the reference: CR+LF line ends, a loop row with more values than
 * labels, or with fewer when there are more than two labels, an unclosed quote. The version tag (# version N) is
 * kept (30000 when absent). */
int rln_star_read(const char *path, rln_star **out);
/* The same from text in memory. */
int rln_star_parse(const char *text, size_t len, rln_star **out);
/* An empty file to fill with rln_star_add_block / _add_column / _add_rows / _set_value. */
int rln_star_create(rln_star **out);
int rln_star_destroy(rln_star *s);

int rln_star_version(const rln_star *s);
int rln_star_block_count(const rln_star *s);
const char *rln_star_block_name(const rln_star *s, int block);
/* Index of the block data_<name>, or -1. */
int rln_star_find_block(const rln_star *s, const char *name);
/* 1 for a label-value list (one row), 0 for a loop_ table. */
int rln_star_block_is_list(const rln_star *s, int block);
long rln_star_rows(const rln_star *s, int block);
int rln_star_columns(const rln_star *s, int block);
/* Label of a column without the leading underscore (e.g. "rlnDefocusU"). */
const char *rln_star_label(const rln_star *s, int block, int column);
/* Index of a label in a block, or -1 (MetaDataTable::containsLabel). */
int rln_star_find_label(const rln_star *s, int block, const char *label);
/* The value's text, quotes removed; NULL when out of range. */
const char *rln_star_value(const rln_star *s, int block, long row, int column);
/* A whole column as numbers (MetaDataTable::getValue for RFLOAT / int labels): RLN_ERR_ARG when a value does not
 * parse completely. */
int rln_star_column_double(const rln_star *s, int block, int column, double *out);
int rln_star_column_long(const rln_star *s, int block, int column, long *out);

int rln_star_add_block(rln_star *s, const char *name, int is_list, int *block);
/* Appends a column whose value in every existing row is default_value ("" is written as ""). */
int rln_star_add_column(rln_star *s, int block, const char *label, const char *default_value, int *column);
/* Appends n rows of default values (a list block holds one row). */
int rln_star_add_rows(rln_star *s, int block, long n);
int rln_star_set_value(rln_star *s, int block, long row, int column, const char *value);
/* MetaDataTable::write for every block: "# version 50001", data_<name>, loop_ with "_label #n" lines and rows of
 * values each right-aligned in 10 characters, or "_label value" lines for lists; values with whitespace or a
 * leading quote are quoted (escapeStringForSTAR, src/strings.cpp:88-118). */
int rln_star_write(const rln_star *s, const char *path);

/* ---------------------------------------------------------------------------------------------
 * MRC images: rwMRC.h (src/rwMRC.h:40-500)
 * ------------------------------------------------------------------------------------------- */
typedef struct rln_mrc_info {
    int nx, ny, nz;          /* columns, rows, sections (images of a stack) */
    int mode;                /* 0 int8, 1 int16, 2 float32, 6 uint16, 12 float16, 101 4-bit (SerialEM) */
    int nsymbt;              /* extended header bytes: data start at 1024 + nsymbt */
    int swapped;             /* 1: the file's byte order is the opposite of this machine's */
    double pixel_size;       /* cell a / mx in Angstrom (0 when the header gives none) */
    double pixel_size_z;     /* cell c / mz */
    float amin, amax, amean, arms;
} rln_mrc_info;

/* parseMRCHeader (src/rwMRC.h:143-245): byte order from |mode| or |nx| above 65535, modes 3 and 4 (complex)
 * refused, mode 101 only with nx * ny even. */
int rln_mrc_read_info(const char *path, rln_mrc_info *info);
/* Sections [first, first + count) as float32 [count][ny][nx] (readData with the mode's conversion; 4-bit pixels
 * low nibble first, castPage2T src/image.h:709-723). */
int rln_mrc_read(const char *path, long first, long count, float *out);
/* writeMRC (src/rwMRC.h:280-500) for float data: mode 2, nz = sections (or images of a stack), mx = nx, cell
 * a = pixel_size * nx (nx when pixel_size <= 0), angles 90, axes 1 2 3, "MAP " and this machine's stamp, amin /
 * amax / amean and arms (standard deviation over N - 1), one label. */
int rln_mrc_write(const char *path, const float *data, int nx, int ny, int nz, double pixel_size);

/* ---------------------------------------------------------------------------------------------
 * Particle images named by a STAR column: "<n>@<stack>" (n from 1) or "<file>" (the first image), as
 * Image::read of rlnImageName entries. Paths are relative to base_dir (NULL or "" for the working directory).
 * Reads rows [first, first + count) of the block into out [count][box][box] (float32), every image box x box;
 * each stack's header is read once, then runs of up to 32 images consecutive in the stack and in the output are
 * read by n_threads threads (0: all cores). Missing stacks, other box sizes and indices beyond a stack are errors. */
int rln_star_read_images(const rln_star *s, int block, int column, long first, long count, const char *base_dir,
                         int box, int n_threads, float *out);

#ifdef __cplusplus
}
#endif

#endif /* RLN_FORMATS_H */
