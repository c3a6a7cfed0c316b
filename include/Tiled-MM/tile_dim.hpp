// Tile extent (rows x cols). Same interface as the reference's gpu::tile_dim
// (Tiled-MM src/Tiled-MM/tile_dim.hpp:4-22); size() is kept consistent with the setters.
#pragma once

namespace gpu {

class tile_dim {
 public:
  tile_dim() = default;
  tile_dim(int n_rows, int n_cols) : rows_(n_rows), cols_(n_cols) {}
  int rows() const { return rows_; }
  int cols() const { return cols_; }
  void set_rows(int n_rows) { rows_ = n_rows; }
  void set_cols(int n_cols) { cols_ = n_cols; }
  int size() const { return rows_ * cols_; }

 private:
  int rows_ = 0;
  int cols_ = 0;
};

}  // namespace gpu
