/*
 * rast_io.h -- host-side asset formats on either side of the rasterizer:
 * Wavefront OBJ meshes and PNG textures in, framebuffers out
 * (librast_io.so, no GPU code; PNG over the system zlib).
 *
 * Reference (kdedeurw/Custom-GPGPU-Rasterizer, CCR = CustomCUDARasterizer/):
 *   ObjParser::OpenFile / ParseData / StorePosition / StoreFace / StoreUV /
 *   StoreNormal / AssignVertices  CCR/ObjParser.cpp:23-584
 *   Elite::Renderer::SaveBackbufferToImage (SDL_SaveBMP)  CCR/ERenderer.cpp:174-177
 *   CUDATexture::Create (IMG_Load + pitched upload)  CCR/CUDATexture.cpp:24-45
 *
 * OBJ parsing follows the reference line by line, including its format
 * assumptions: a position line is "v" followed by two spaces (the 3ds Max
 * exporter's layout; the first three characters are skipped), "vt", "vn" and
 * "f" lines by one space, fields separated by single spaces, only the first
 * three vertices of a face, 1-based indices.  Vertices are de-duplicated by
 * (position index, uv index, normal value with a 2-ulp tolerance) and, when
 * the file has both uvs and normals, per-vertex tangents are accumulated
 * from the uv gradients and made orthogonal to the normal
 * (CCR/ObjParser.cpp:446-473).  Deviations (the reference reads memory it
 * never wrote): tangents start from zero, the colour of every vertex is
 * (1, 1, 1), normals are zero when the file has none.  Input the reference
 * cannot parse (an empty index field as in "1//3", negative or out-of-range
 * indices, a face with fewer than three vertices) is rejected with an error.
 */
#ifndef RAST_IO_H
#define RAST_IO_H

#include <stddef.h>
#include <stdint.h>

#include "rast.h"

#ifdef __cplusplus
extern "C" {
#endif

/* VertexType flags of CCR/VertexType.h:3-11. */
enum { RIO_VERTEX_UV = 1, RIO_VERTEX_NORMAL = 4, RIO_VERTEX_TANGENT = 8 };

/* ObjParser::ReadFromObjFile(std::vector<IVertex>&, std::vector<unsigned>&,
 * short&) (CCR/ObjParser.cpp:100-109, 381-480) on the text of a file.  invert_y =
 * SetInvertYAxis (v -> 1 - v).  On success *verts / *indices are arrays the
 * caller releases with rio_free; *vertex_type gets RIO_VERTEX_* flags. */
int rio_obj_parse(const char *text, size_t len, int invert_y, rast_ivertex **verts, uint32_t *num_verts,
                  uint32_t **indices, uint32_t *num_indices, int *vertex_type);
/* OpenFile (CCR/ObjParser.cpp:564-584: the extension must be ".obj") + parse. */
int rio_obj_load(const char *path, int invert_y, rast_ivertex **verts, uint32_t *num_verts, uint32_t **indices,
                 uint32_t *num_indices, int *vertex_type);
void rio_free(void *p);

/* The back buffer to a 32-bit uncompressed BMP (bottom-up rows, B,G,R,A
 * bytes), as SDL_SaveBMP writes a 32-bit surface; rgba8 texels in memory
 * order R,G,B,A, rows of width texels. */
int rio_write_bmp(const char *path, const uint32_t *rgba8, uint32_t width, uint32_t height);

/* IMG_Load for the reference's textures (CCR/CUDATexture.cpp:24-45): a PNG
 * file's bytes decoded to rgba8 texels (memory order R,G,B,A, rows top to
 * bottom, the layout rast_texture_create takes).  Bit depth 8; grey, RGB,
 * palette, grey + alpha and RGBA; not interlaced; every chunk's CRC is
 * checked.  *texels is released with rio_free. */
int rio_png_decode(const uint8_t *data, size_t len, uint32_t **texels, uint32_t *width, uint32_t *height);
/* The same from a file. */
int rio_png_load(const char *path, uint32_t **texels, uint32_t *width, uint32_t *height);
/* rgba8 texels to an 8-bit RGBA PNG (no counterpart in the reference: test
 * data and frame output).  filter 0..4 filters every row with that type,
 * -1 row y with type y % 5.  *out is released with rio_free. */
int rio_png_encode(const uint32_t *rgba8, uint32_t width, uint32_t height, int filter, uint8_t **out,
                   size_t *out_len);

/* Message of the calling thread's last failure. */
int rio_last_error(char *buf, size_t n);

#ifdef __cplusplus
}
#endif
#endif
