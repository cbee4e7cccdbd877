"""pyemma.coordinates-compatible entry points of the GPU discretization path."""
from .api import assign_to_centers, cluster_kmeans, covariance_lagged, tica
from .clustering import AssignCenters, KmeansClustering
from .covariance import LaggedCovariance
from .tica import TICA

__all__ = ["tica", "covariance_lagged", "cluster_kmeans", "assign_to_centers", "TICA", "LaggedCovariance",
           "KmeansClustering", "AssignCenters"]
